! lesgo_amd_iface.f90 -- ISO_C_BINDING interface to liblesgo_hip.so (include/lesgo_amd.h)
! for a Fortran host (the reference's own language).  Derived types mirror the C
! structures byte for byte; every routine keeps the reference's argument meaning.
module lesgo_amd
  use iso_c_binding
  implicit none
  private
  public :: lp_grid, lp_droplet_config, lp_config
  public :: lp_grid_init, lp_config_default, lp_create, lp_destroy, lp_sync, lp_last_error_f
  public :: lp_upload, lp_download, lp_field_info, lp_set_time, lp_get_time
  public :: lp_init_wavenumber, lp_test_filter_init, lp_filt_da, lp_ddx, lp_ddy, lp_ddxy
  public :: lp_ddz_uv, lp_ddz_w, lp_padd, lp_unpadd, lp_dealias1, lp_dealias2, lp_test_filter
  public :: lp_divstress_uv, lp_divstress_w, lp_tridag_array
  public :: lp_wallstress, lp_sgs_stag, lp_lagrange_sdep, lp_interpolag_sdep, lp_convec
  public :: lp_press_stag_array, lp_project, lp_timestep_conditions, lp_rmsdiv, lp_step
  public :: lp_plane_stats, lp_spectrum, lp_energy
  public :: lp_tavg_reset, lp_tavg_compute, lp_tavg_get
  public :: lp_pollen_all_in_one2, lp_calculate_probability
  public :: lp_pollen_rhs_calc2, lp_calcbeta_pcon, lp_step_pollen2
  public :: lp_set_cfl_limit, lp_check_finite, lp_get_precision
  public :: LP_OK, LP_FP64, LP_MIXED, LP_LBC_WALL, LP_LBC_STRESS_FREE
  public :: LP_U, LP_V, LP_W, LP_P, LP_RHSX, LP_RHSY, LP_RHSZ
  public :: LP_PCON, LP_RHS_PCON, LP_RHS_PCON_F, LP_RHS_PBE, LP_BETA_PCON
  public :: LP_THETA, LP_RHS_T, LP_RHS_TF, LP_SGS_T3, LP_BETA_SCAL, LP_PSI_M
  public :: lp_scalar_config, lp_obukhov, lp_obukhov_get, lp_theta_all_in_one

  integer(c_int), parameter :: LP_OK = 0
  integer(c_int32_t), parameter :: LP_FP64 = 0, LP_MIXED = 1
  integer(c_int32_t), parameter :: LP_LBC_WALL = 0, LP_LBC_STRESS_FREE = 1
  ! field ids (enum lp_field)
  integer(c_int32_t), parameter :: LP_U = 0, LP_V = 1, LP_W = 2
  integer(c_int32_t), parameter :: LP_RHSX = 12, LP_RHSY = 13, LP_RHSZ = 14, LP_P = 27
  integer(c_int32_t), parameter :: LP_PCON = 47, LP_RHS_PCON = 48, LP_RHS_PCON_F = 49
  integer(c_int32_t), parameter :: LP_RHS_PBE = 50, LP_BETA_PCON = 51
  integer(c_int32_t), parameter :: LP_THETA = 52, LP_RHS_T = 53, LP_RHS_TF = 54, LP_SGS_T3 = 58
  integer(c_int32_t), parameter :: LP_BETA_SCAL = 59, LP_PSI_M = 60

  type, bind(C) :: lp_grid
    integer(c_int32_t) :: nx, ny, nz, ld, lh, nx2, ny2, lh_big, ld_big, pad_
    real(c_double) :: L_x, L_y, L_z, dx, dy, dz
  end type lp_grid

  type, bind(C) :: lp_droplet_config
    integer(c_int32_t) :: npcon, pbe, pbe_feedback, settling, inertia, active
    integer(c_int32_t) :: bc_x, bc_y, src_i, src_j, src_k, src_ngau, src_jt0, src_jt1
    integer(c_int32_t) :: pcon_init, init_pcon
    real(c_double) :: src_sigma, src_frac, Sc, sfc_flux, rho_c, rho_w
    real(c_double) :: sigma_s, nu_c, nu_d, k_b, d_min, g, u_star, z_i
    type(c_ptr) :: diameter, settling_vel, Q_src
  end type lp_droplet_config

  type, bind(C) :: lp_scalar_config
    integer(c_int32_t) :: s_flag, passive, scal_init, initsc, lbc, ocean, iboussinesq, pad_
    real(c_double) :: Pr, wt_s, T_scale, dTdz_top, g, u_star, z_i
  end type lp_scalar_config

  type, bind(C) :: lp_config
    type(lp_grid) :: grid
    real(c_double) :: dt
    integer(c_int32_t) :: precision, model, inilag, initu, cs_count, dyn_init, fp32_casts, lbc_mom
    real(c_double) :: z0, d0, vonk, Co
    integer(c_int32_t) :: nnn, use_mean_p_force
    real(c_double) :: mean_p_force, cs2_init, opftime, filter_size
    type(lp_droplet_config) :: drop
    type(lp_scalar_config) :: scal
  end type lp_config

  interface
    integer(c_int) function lp_grid_init(g, nx, ny, nz, L_x, L_y, L_z) bind(C, name="lp_grid_init")
      import :: c_int, c_int32_t, c_double, lp_grid
      type(lp_grid), intent(inout) :: g
      integer(c_int32_t), value :: nx, ny, nz
      real(c_double), value :: L_x, L_y, L_z
    end function
    subroutine lp_config_default(c) bind(C, name="lp_config_default")
      import :: lp_config
      type(lp_config), intent(inout) :: c
    end subroutine
    integer(c_int) function lp_create(cfg, ctx) bind(C, name="lp_create")
      import :: c_int, c_ptr, lp_config
      type(lp_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function lp_destroy(ctx) bind(C, name="lp_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_sync(ctx) bind(C, name="lp_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function lp_last_error() bind(C, name="lp_last_error")
      import :: c_ptr
    end function
    integer(c_int) function lp_field_info(ctx, field, dev_ptr, bytes, esz) bind(C, name="lp_field_info")
      import :: c_int, c_ptr, c_int32_t, c_size_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: field
      type(c_ptr), intent(out) :: dev_ptr
      integer(c_size_t), intent(out) :: bytes
      integer(c_int32_t), intent(out) :: esz
    end function
    integer(c_int) function lp_upload(ctx, field, host, bytes) bind(C, name="lp_upload")
      import :: c_int, c_ptr, c_int32_t, c_size_t
      type(c_ptr), value :: ctx, host
      integer(c_int32_t), value :: field
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function lp_download(ctx, field, host, bytes) bind(C, name="lp_download")
      import :: c_int, c_ptr, c_int32_t, c_size_t
      type(c_ptr), value :: ctx, host
      integer(c_int32_t), value :: field
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function lp_set_time(ctx, jt, jt_total) bind(C, name="lp_set_time")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: jt, jt_total
    end function
    integer(c_int) function lp_get_time(ctx, jt, jt_total) bind(C, name="lp_get_time")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: jt, jt_total
    end function
    ! ---- spectral primitives: arrays passed by reference (first element) ----
    integer(c_int) function lp_init_wavenumber(ctx, kx, ky, k2) bind(C, name="lp_init_wavenumber")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: kx(*), ky(*), k2(*)
    end function
    integer(c_int) function lp_test_filter_init(ctx, alpha, G) bind(C, name="lp_test_filter_init")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: alpha
      real(c_double) :: G(*)
    end function
    integer(c_int) function lp_filt_da(ctx, f, dfdx, dfdy, nlev) bind(C, name="lp_filt_da")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: f(*), dfdx(*), dfdy(*)
      integer(c_int32_t), value :: nlev
    end function
    integer(c_int) function lp_ddx(ctx, dfdx, f, nlev) bind(C, name="lp_ddx")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: dfdx(*), f(*)
      integer(c_int32_t), value :: nlev
    end function
    integer(c_int) function lp_ddy(ctx, dfdy, f, nlev) bind(C, name="lp_ddy")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: dfdy(*), f(*)
      integer(c_int32_t), value :: nlev
    end function
    integer(c_int) function lp_ddxy(ctx, dfdx, dfdy, f, nlev) bind(C, name="lp_ddxy")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: dfdx(*), dfdy(*), f(*)
      integer(c_int32_t), value :: nlev
    end function
    integer(c_int) function lp_ddz_uv(ctx, dfdz, f) bind(C, name="lp_ddz_uv")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: dfdz(*), f(*)
    end function
    integer(c_int) function lp_ddz_w(ctx, dfdz, f) bind(C, name="lp_ddz_w")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: dfdz(*), f(*)
    end function
    integer(c_int) function lp_padd(ctx, u_big, u, nplanes) bind(C, name="lp_padd")
      import :: c_int, c_ptr, c_int32_t, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: u_big(*), u(*)
      integer(c_int32_t), value :: nplanes
    end function
    integer(c_int) function lp_unpadd(ctx, cc, cc_big, nplanes) bind(C, name="lp_unpadd")
      import :: c_int, c_ptr, c_int32_t, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: cc(*), cc_big(*)
      integer(c_int32_t), value :: nplanes
    end function
    integer(c_int) function lp_dealias1(ctx, u, u_big, nlev) bind(C, name="lp_dealias1")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), u_big(*)
      integer(c_int32_t), value :: nlev
    end function
    integer(c_int) function lp_dealias2(ctx, u, u_big, nlev) bind(C, name="lp_dealias2")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), u_big(*)
      integer(c_int32_t), value :: nlev
    end function
    integer(c_int) function lp_test_filter(ctx, f, G, nplanes) bind(C, name="lp_test_filter")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: f(*), G(*)
      integer(c_int32_t), value :: nplanes
    end function
    integer(c_int) function lp_divstress_uv(ctx, divt, tx, ty, tz) bind(C, name="lp_divstress_uv")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: divt(*), tx(*), ty(*), tz(*)
    end function
    integer(c_int) function lp_divstress_w(ctx, divt, tx, ty, tz) bind(C, name="lp_divstress_w")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: divt(*), tx(*), ty(*), tz(*)
    end function
    integer(c_int) function lp_tridag_array(ctx, a, b, c, r, u) bind(C, name="lp_tridag_array")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double) :: a(*), b(*), c(*)
      complex(c_double_complex) :: r(*), u(*)
    end function
    ! ---- routines on the context state (the reference's no-argument routines) ----
    integer(c_int) function lp_wallstress(ctx) bind(C, name="lp_wallstress")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_sgs_stag(ctx) bind(C, name="lp_sgs_stag")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_lagrange_sdep(ctx) bind(C, name="lp_lagrange_sdep")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_interpolag_sdep(ctx) bind(C, name="lp_interpolag_sdep")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_convec(ctx) bind(C, name="lp_convec")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_press_stag_array(ctx) bind(C, name="lp_press_stag_array")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_project(ctx) bind(C, name="lp_project")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_timestep_conditions(ctx, cfl, visc) bind(C, name="lp_timestep_conditions")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: cfl, visc
    end function
    integer(c_int) function lp_rmsdiv(ctx, rms) bind(C, name="lp_rmsdiv")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rms
    end function
    ! (10, nz-1) in Fortran order: the per-level statistics of avg_stats (io.f90:1782-1849)
    integer(c_int) function lp_plane_stats(ctx, stats) bind(C, name="lp_plane_stats")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: stats(*)
    end function
    ! (nx/2, nz-1) in Fortran order: spectrum (io.f90:1643-1679) of every level of a field
    integer(c_int) function lp_spectrum(ctx, field, bin, spec) bind(C, name="lp_spectrum")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: field, bin
      real(c_double), intent(out) :: spec(*)
    end function
    ! obukhov, theta_all_in_one (scalars_module_crossflow_mod.f90:141-360, 1170-1403)
    integer(c_int) function lp_obukhov(ctx) bind(C, name="lp_obukhov")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_obukhov_get(ctx, vals) bind(C, name="lp_obukhov_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: vals(8)
    end function
    integer(c_int) function lp_theta_all_in_one(ctx) bind(C, name="lp_theta_all_in_one")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    ! energy (energy.f90:1-47): kea
    integer(c_int) function lp_energy(ctx, kea) bind(C, name="lp_energy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: kea
    end function
    ! tavg_init, tavg_compute (io_crossflow_mod.f90:2452-2799)
    integer(c_int) function lp_tavg_reset(ctx) bind(C, name="lp_tavg_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_tavg_compute(ctx, tavg_dt) bind(C, name="lp_tavg_compute")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: tavg_dt
    end function
    ! accumulator q (0-based, the order of the tavg struct) into out(ld, ny, nz)
    integer(c_int) function lp_tavg_get(ctx, q, out, nbytes, total_time) bind(C, name="lp_tavg_get")
      import :: c_int, c_int32_t, c_ptr, c_double, c_size_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: q
      type(c_ptr), value :: out
      integer(c_size_t), value :: nbytes
      real(c_double), intent(out) :: total_time
    end function
    integer(c_int) function lp_pollen_all_in_one2(ctx) bind(C, name="lp_pollen_all_in_one2")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_pollen_rhs_calc2(ctx) bind(C, name="lp_pollen_rhs_calc2")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_calcbeta_pcon(ctx) bind(C, name="lp_calcbeta_pcon")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_step_pollen2(ctx) bind(C, name="lp_step_pollen2")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lp_calculate_probability(diameter, npcon, sigma_s, d_min, beta) &
        bind(C, name="lp_calculate_probability")
      import :: c_int, c_int32_t, c_double
      real(c_double), intent(in) :: diameter(*)
      integer(c_int32_t), value :: npcon
      real(c_double), value :: sigma_s, d_min
      real(c_double), intent(out) :: beta(*)
    end function
    integer(c_int) function lp_step(ctx, nsteps) bind(C, name="lp_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function lp_set_cfl_limit(ctx, cfl_max) bind(C, name="lp_set_cfl_limit")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cfl_max
    end function
    integer(c_int) function lp_check_finite(ctx) bind(C, name="lp_check_finite")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    ! the explicit-array routines above take binary64 arrays: an LP_FP64 context
    integer(c_int) function lp_get_precision(ctx, precision) bind(C, name="lp_get_precision")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: precision
    end function
  end interface

contains

  ! lp_last_error() as a Fortran string
  function lp_last_error_f() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: i
    msg = ' '
    cp = lp_last_error()
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function lp_last_error_f

end module lesgo_amd
