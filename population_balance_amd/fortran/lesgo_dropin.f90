! lesgo_dropin.f90 -- the reference's spectral and stress primitives under their own
! names and argument lists, as external subroutines with implicit interfaces (the way
! the reference declares and calls them, e.g. `call ddx(dtxdx, tx)`), running on the
! GPU through liblesgo_hip.so.  Linking this object in place of the reference's
! filt_da.f90 ... tridag_array.f90 objects switches those call sites without editing
! them.  Extents come from the context's grid, as the reference takes them from param.
!
!   filt_da(f, dfdx, dfdy)            filt_da.f90:3-36       (ld, ny, nz) x 3
!   ddx(dfdx, f), ddy(dfdy, f)        ddx.f90:1-31, ddy.f90:1-31
!   ddxy(dfdx, dfdy, f)               ddxy.f90:1-29
!   ddz_uv(dfdz, f), ddz_w(dfdz, f)   ddz_uv.f90:5-60, ddz_w.f90:6-51
!   padd(u_big, u), unpadd(cc, cc_big)   padd.f90:1-26, unpadd.f90:1-22 (one plane)
!   dealias1(u, u_big), dealias2(u, u_big)   dealias1.f90:1-31, dealias2.f90:1-33
!   divstress_uv / divstress_w(divt, tx, ty, tz)   divstress_uv.f90:2-129, divstress_w.f90:4-85
!   tridag_array(a, b, c, r, u)       tridag_array.f90:1-163
!
! The reference stops on errors (tridag_array.f90:35-38,70-75); so do these, with the
! library's message.  The arrays are binary64 (rprec), so the context must be LP_FP64:
! lesgo_dropin_init refuses any other.
module lesgo_dropin
  use iso_c_binding
  use lesgo_amd
  implicit none
  private
  public :: lesgo_dropin_init, lesgo_dropin_finalize, dropin_check, dropin_ctx, dropin_nz
  type(c_ptr), save :: dropin_ctx = c_null_ptr
  integer(c_int32_t), save :: dropin_nz = 0
contains
  ! create the context the shims use (plans, wavenumbers, filters: fft.f90:35-76)
  subroutine lesgo_dropin_init(cfg)
    type(lp_config), intent(in) :: cfg
    integer(c_int32_t) :: prec
    if (cfg%precision /= LP_FP64) then
      write (*, '(a)') 'lesgo_dropin_init: the reference-signature routines take real(rprec) arrays; '// &
                       'create the context with precision = LP_FP64'
      error stop 1
    end if
    call dropin_check(lp_create(cfg, dropin_ctx), 'lp_create')
    call dropin_check(lp_get_precision(dropin_ctx, prec), 'lp_get_precision')
    if (prec /= LP_FP64) error stop 1
    dropin_nz = cfg%grid%nz
  end subroutine lesgo_dropin_init

  subroutine lesgo_dropin_finalize()
    integer(c_int) :: rc
    if (c_associated(dropin_ctx)) rc = lp_destroy(dropin_ctx)
    dropin_ctx = c_null_ptr
  end subroutine lesgo_dropin_finalize

  subroutine dropin_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= LP_OK) then
      write (*, '(a,a,i0,a,a)') what, ' failed (', rc, '): ', trim(lp_last_error_f())
      error stop 2
    end if
  end subroutine dropin_check
end module lesgo_dropin

subroutine filt_da(f, dfdx, dfdy)
  use lesgo_dropin
  use lesgo_amd, only: lp_filt_da
  use iso_c_binding
  implicit none
  real(c_double) :: f(*), dfdx(*), dfdy(*)
  call dropin_check(lp_filt_da(dropin_ctx, f, dfdx, dfdy, dropin_nz), 'filt_da')
end subroutine filt_da

subroutine ddx(dfdx, f)
  use lesgo_dropin
  use lesgo_amd, only: lp_ddx
  use iso_c_binding
  implicit none
  real(c_double) :: dfdx(*), f(*)
  call dropin_check(lp_ddx(dropin_ctx, dfdx, f, dropin_nz), 'ddx')
end subroutine ddx

subroutine ddy(dfdy, f)
  use lesgo_dropin
  use lesgo_amd, only: lp_ddy
  use iso_c_binding
  implicit none
  real(c_double) :: dfdy(*), f(*)
  call dropin_check(lp_ddy(dropin_ctx, dfdy, f, dropin_nz), 'ddy')
end subroutine ddy

subroutine ddxy(dfdx, dfdy, f)
  use lesgo_dropin
  use lesgo_amd, only: lp_ddxy
  use iso_c_binding
  implicit none
  real(c_double) :: dfdx(*), dfdy(*), f(*)
  call dropin_check(lp_ddxy(dropin_ctx, dfdx, dfdy, f, dropin_nz), 'ddxy')
end subroutine ddxy

subroutine ddz_uv(dfdz, f)
  use lesgo_dropin
  use lesgo_amd, only: lp_ddz_uv
  use iso_c_binding
  implicit none
  real(c_double) :: dfdz(*), f(*)
  call dropin_check(lp_ddz_uv(dropin_ctx, dfdz, f), 'ddz_uv')
end subroutine ddz_uv

subroutine ddz_w(dfdz, f)
  use lesgo_dropin
  use lesgo_amd, only: lp_ddz_w
  use iso_c_binding
  implicit none
  real(c_double) :: dfdz(*), f(*)
  call dropin_check(lp_ddz_w(dropin_ctx, dfdz, f), 'ddz_w')
end subroutine ddz_w

subroutine padd(u_big, u)
  use lesgo_dropin
  use lesgo_amd, only: lp_padd
  use iso_c_binding
  implicit none
  complex(c_double_complex) :: u_big(*), u(*)
  call dropin_check(lp_padd(dropin_ctx, u_big, u, 1_c_int32_t), 'padd')
end subroutine padd

subroutine unpadd(cc, cc_big)
  use lesgo_dropin
  use lesgo_amd, only: lp_unpadd
  use iso_c_binding
  implicit none
  complex(c_double_complex) :: cc(*), cc_big(*)
  call dropin_check(lp_unpadd(dropin_ctx, cc, cc_big, 1_c_int32_t), 'unpadd')
end subroutine unpadd

subroutine dealias1(u, u_big)
  use lesgo_dropin
  use lesgo_amd, only: lp_dealias1
  use iso_c_binding
  implicit none
  real(c_double) :: u(*), u_big(*)
  call dropin_check(lp_dealias1(dropin_ctx, u, u_big, dropin_nz), 'dealias1')
end subroutine dealias1

subroutine dealias2(u, u_big)
  use lesgo_dropin
  use lesgo_amd, only: lp_dealias2
  use iso_c_binding
  implicit none
  real(c_double) :: u(*), u_big(*)
  call dropin_check(lp_dealias2(dropin_ctx, u, u_big, dropin_nz), 'dealias2')
end subroutine dealias2

subroutine divstress_uv(divt, tx, ty, tz)
  use lesgo_dropin
  use lesgo_amd, only: lp_divstress_uv
  use iso_c_binding
  implicit none
  real(c_double) :: divt(*), tx(*), ty(*), tz(*)
  call dropin_check(lp_divstress_uv(dropin_ctx, divt, tx, ty, tz), 'divstress_uv')
end subroutine divstress_uv

subroutine divstress_w(divt, tx, ty, tz)
  use lesgo_dropin
  use lesgo_amd, only: lp_divstress_w
  use iso_c_binding
  implicit none
  real(c_double) :: divt(*), tx(*), ty(*), tz(*)
  call dropin_check(lp_divstress_w(dropin_ctx, divt, tx, ty, tz), 'divstress_w')
end subroutine divstress_w

subroutine tridag_array(a, b, c, r, u)
  use lesgo_dropin
  use lesgo_amd, only: lp_tridag_array
  use iso_c_binding
  implicit none
  real(c_double) :: a(*), b(*), c(*)
  complex(c_double_complex) :: r(*), u(*)
  call dropin_check(lp_tridag_array(dropin_ctx, a, b, c, r, u), 'tridag_array')
end subroutine tridag_array
