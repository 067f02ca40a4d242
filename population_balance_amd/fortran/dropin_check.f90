! dropin_check.f90 -- a Fortran host that calls the reference's primitives by their own
! names through implicit interfaces (as the reference's call sites do), with arrays
! declared as the reference declares them, on fields whose results are known in closed
! form.  Linked against lesgo_dropin.f90, so every call runs on the GPU.  Stops with a
! non-zero code at the first result outside 1e-12 (relative to the field's scale).
program dropin_check
  use iso_c_binding
  use lesgo_amd
  use lesgo_dropin
  implicit none
  integer, parameter :: rprec = c_double
  real(rprec), parameter :: bogus = -1234567890.0_rprec
  external :: filt_da, ddx, ddy, ddxy, ddz_uv, ddz_w, padd, unpadd, dealias1, dealias2
  external :: divstress_uv, divstress_w, tridag_array
  type(lp_config), target :: cfg
  integer :: nx, ny, nz, ld, lh, nx2, ny2, ld_big, lh_big, jx, jy, jz, m
  real(rprec), allocatable :: f(:,:,:), dfdx(:,:,:), dfdy(:,:,:), dfdz(:,:,:), g(:,:,:)
  real(rprec), allocatable :: ub(:,:,:), tx(:,:,:), ty(:,:,:), tz(:,:,:), divt(:,:,:)
  real(rprec), allocatable :: a(:,:,:), b(:,:,:), c(:,:,:)
  complex(rprec), allocatable :: r(:,:,:), u(:,:,:), ue(:,:,:)
  complex(rprec), allocatable :: sp(:,:), sb(:,:), sc(:,:)
  real(rprec) :: x, y, dx, dy, dz, err, e, pi
  integer(c_int32_t) :: rc32

  call lp_config_default(cfg)
  pi = 4.0_rprec*atan(1.0_rprec)
  rc32 = lp_grid_init(cfg%grid, 32_c_int32_t, 32_c_int32_t, 9_c_int32_t, 2*pi, 2*pi, 1.0_rprec)
  call lesgo_dropin_init(cfg)
  nx = cfg%grid%nx; ny = cfg%grid%ny; nz = cfg%grid%nz; ld = cfg%grid%ld; lh = cfg%grid%lh
  nx2 = cfg%grid%nx2; ny2 = cfg%grid%ny2; ld_big = cfg%grid%ld_big; lh_big = cfg%grid%lh_big
  dx = cfg%grid%dx; dy = cfg%grid%dy; dz = cfg%grid%dz
  allocate (f(ld, ny, nz), dfdx(ld, ny, nz), dfdy(ld, ny, nz), dfdz(ld, ny, nz), g(ld, ny, nz))
  allocate (ub(ld_big, ny2, nz), tx(ld, ny, nz), ty(ld, ny, nz), tz(ld, ny, nz), divt(ld, ny, nz))

  ! ---- filt_da: the Nyquist waves cos(16x), cos(16y) are removed, the rest is kept
  f = 0
  do jz = 1, nz
    do jy = 1, ny
      do jx = 1, nx
        x = (jx - 1)*dx; y = (jy - 1)*dy
        f(jx, jy, jz) = sin(3*x)*cos(5*y) + jz + 0.5_rprec*cos(16*x) + 0.25_rprec*cos(16*y)
      end do
    end do
  end do
  call filt_da(f, dfdx, dfdy)
  err = 0
  do jz = 1, nz
    do jy = 1, ny
      do jx = 1, nx
        x = (jx - 1)*dx; y = (jy - 1)*dy
        err = max(err, abs(f(jx, jy, jz) - (sin(3*x)*cos(5*y) + jz))/nz)
        err = max(err, abs(dfdx(jx, jy, jz) - 3*cos(3*x)*cos(5*y))/3)
        err = max(err, abs(dfdy(jx, jy, jz) + 5*sin(3*x)*sin(5*y))/5)
      end do
    end do
  end do
  call report('filt_da', err, 1)

  ! ---- ddx, ddy, ddxy on sin(3x) cos(5y)
  call fill_wave(f)
  dfdx = 0; dfdy = 0
  call ddx(dfdx, f)
  call ddy(dfdy, f)
  err = wave_err(dfdx, dfdy)
  call report('ddx / ddy', err, 2)
  dfdx = 0; dfdy = 0
  call ddxy(dfdx, dfdy, f)
  call report('ddxy', wave_err(dfdx, dfdy), 3)

  ! ---- ddz_uv of z^2 on uvp heights: 2 z on w-nodes (level 1 untouched, top 0);
  !      ddz_w of z^2 on w heights: 2 z on uvp nodes (top 0)
  do jz = 1, nz
    f(:, :, jz) = ((jz - 0.5_rprec)*dz)**2
    g(:, :, jz) = ((jz - 1)*dz)**2
  end do
  dfdz = 7.0_rprec
  call ddz_uv(dfdz, f)
  err = maxval(abs(dfdz(1:nx, :, 1) - 7.0_rprec)) + maxval(abs(dfdz(1:nx, :, nz)))
  do jz = 2, nz - 1
    err = max(err, maxval(abs(dfdz(1:nx, :, jz) - 2*(jz - 1)*dz)))
  end do
  call report('ddz_uv', err, 4)
  call ddz_w(dfdz, g)
  err = maxval(abs(dfdz(1:nx, :, nz)))
  do jz = 1, nz - 1
    err = max(err, maxval(abs(dfdz(1:nx, :, jz) - 2*(jz - 0.5_rprec)*dz)))
  end do
  call report('ddz_w', err, 5)

  ! ---- padd / unpadd on one plane (padd.f90, unpadd.f90)
  allocate (sp(lh, ny), sb(lh_big, ny2), sc(lh, ny))
  do jy = 1, ny
    do jx = 1, lh
      sp(jx, jy) = cmplx(jx + 0.1_rprec*jy, jx - 0.3_rprec*jy, rprec)
    end do
  end do
  sb = (99.0_rprec, 99.0_rprec)
  call padd(sb, sp)
  err = 0
  do jy = 1, ny2
    do jx = 1, lh_big
      if (jx <= nx/2 .and. jy <= ny/2) then
        err = max(err, abs(sb(jx, jy) - sp(jx, jy)))
      else if (jx <= nx/2 .and. jy >= ny2 - ny/2 + 2) then
        err = max(err, abs(sb(jx, jy) - sp(jx, jy - ny2 + ny)))
      else
        err = max(err, abs(sb(jx, jy)))
      end if
    end do
  end do
  call report('padd', err, 6)
  sc = (55.0_rprec, 55.0_rprec)
  call unpadd(sc, sb)
  err = 0
  do jy = 1, ny
    do jx = 1, lh
      if (jx == lh .or. jy == ny/2 + 1) then
        err = max(err, abs(sc(jx, jy)))
      else
        err = max(err, abs(sc(jx, jy) - sp(jx, jy)))
      end if
    end do
  end do
  call report('unpadd', err, 7)

  ! ---- dealias1: spectral interpolation onto the 3/2 grid; dealias2: truncation back
  !      (the wave cos(20x) cos(2y) of the 3/2 grid lies beyond nx/2 and is removed)
  do jz = 1, nz
    do jy = 1, ny
      do jx = 1, nx
        x = (jx - 1)*dx; y = (jy - 1)*dy
        f(jx, jy, jz) = sin(3*x)*cos(5*y) + jz
      end do
    end do
  end do
  ub = 0
  call dealias1(f, ub)
  err = 0
  do jz = 1, nz
    do jy = 1, ny2
      do jx = 1, nx2
        x = (jx - 1)*2*pi/nx2; y = (jy - 1)*2*pi/ny2
        err = max(err, abs(ub(jx, jy, jz) - (sin(3*x)*cos(5*y) + jz))/nz)
      end do
    end do
  end do
  call report('dealias1', err, 8)
  do jz = 1, nz
    do jy = 1, ny2
      do jx = 1, nx2
        x = (jx - 1)*2*pi/nx2; y = (jy - 1)*2*pi/ny2
        ub(jx, jy, jz) = sin(3*x)*cos(5*y) + cos(20*x)*cos(2*y)
      end do
    end do
  end do
  g = 0
  call dealias2(g, ub)
  err = 0
  do jz = 1, nz
    do jy = 1, ny
      do jx = 1, nx
        x = (jx - 1)*dx; y = (jy - 1)*dy
        err = max(err, abs(g(jx, jy, jz) - sin(3*x)*cos(5*y)))
      end do
    end do
  end do
  call report('dealias2', err, 9)

  ! ---- divstress_uv / divstress_w: d tx/dx + d ty/dy + staggered d tz/dz
  tx = 0; ty = 0; tz = 0
  do jz = 1, nz
    do jy = 1, ny
      do jx = 1, nx
        x = (jx - 1)*dx; y = (jy - 1)*dy
        tx(jx, jy, jz) = sin(2*x)*(1 + 0.1_rprec*jz)
        ty(jx, jy, jz) = cos(3*y)*(2 - 0.2_rprec*jz)
        tz(jx, jy, jz) = 0.3_rprec*jz*jz + sin(x)*0
      end do
    end do
  end do
  call divstress_uv(divt, tx, ty, tz)
  err = maxval(abs(divt(1:nx, :, nz) - bogus))
  do jz = 1, nz - 1
    do jy = 1, ny
      do jx = 1, nx
        x = (jx - 1)*dx; y = (jy - 1)*dy
        e = 2*cos(2*x)*(1 + 0.1_rprec*jz) - 3*sin(3*y)*(2 - 0.2_rprec*jz) + 0.3_rprec*((jz + 1)**2 - jz**2)/dz
        err = max(err, abs(divt(jx, jy, jz) - e)/(1/dz))
      end do
    end do
  end do
  call report('divstress_uv', err, 10)
  call divstress_w(divt, tx, ty, tz)
  err = 0
  do jz = 1, nz
    do jy = 1, ny
      do jx = 1, nx
        x = (jx - 1)*dx; y = (jy - 1)*dy
        e = 2*cos(2*x)*(1 + 0.1_rprec*jz) - 3*sin(3*y)*(2 - 0.2_rprec*jz)
        if (jz >= 2 .and. jz <= nz - 1) e = e + 0.3_rprec*(jz**2 - (jz - 1)**2)/dz
        err = max(err, abs(divt(jx, jy, jz) - e)/(1/dz))
      end do
    end do
  end do
  call report('divstress_w', err, 11)

  ! ---- tridag_array: diagonally dominant systems per mode with a known solution
  allocate (a(lh, ny, nz + 1), b(lh, ny, nz + 1), c(lh, ny, nz + 1))
  allocate (r(lh, ny, nz + 1), u(lh, ny, nz + 1), ue(lh, ny, nz + 1))
  do m = 1, nz + 1
    do jy = 1, ny
      do jx = 1, lh
        a(jx, jy, m) = 1.0_rprec + 0.01_rprec*jx
        c(jx, jy, m) = 0.5_rprec - 0.02_rprec*jy
        b(jx, jy, m) = -(4.0_rprec + 0.1_rprec*(jx + jy) + 0.05_rprec*m)
        ue(jx, jy, m) = cmplx(jx + 0.5_rprec*m, jy - 0.25_rprec*m, rprec)
      end do
    end do
  end do
  do m = 1, nz + 1
    r(:, :, m) = b(:, :, m)*ue(:, :, m)
    if (m > 1) r(:, :, m) = r(:, :, m) + a(:, :, m)*ue(:, :, m - 1)
    if (m < nz + 1) r(:, :, m) = r(:, :, m) + c(:, :, m)*ue(:, :, m + 1)
  end do
  u = 0
  call tridag_array(a, b, c, r, u)
  err = 0
  do jy = 1, ny
    do jx = 1, lh
      if (jx == lh .or. jy == ny/2 + 1 .or. (jx == 1 .and. jy == 1)) cycle  ! skipped modes
      err = max(err, maxval(abs(u(jx, jy, :) - ue(jx, jy, :)))/maxval(abs(ue(jx, jy, :))))
    end do
  end do
  call report('tridag_array', err, 12)

  call lesgo_dropin_finalize()
  write (*, '(a)') 'drop-in check ok'

contains

  subroutine fill_wave(h)
    real(rprec), intent(out) :: h(:,:,:)
    integer :: i, j
    h = 0
    do j = 1, ny
      do i = 1, nx
        h(i, j, :) = sin(3*(i - 1)*dx)*cos(5*(j - 1)*dy)
      end do
    end do
  end subroutine fill_wave

  real(rprec) function wave_err(hx, hy)
    real(rprec), intent(in) :: hx(:,:,:), hy(:,:,:)
    integer :: i, j
    wave_err = 0
    do j = 1, ny
      do i = 1, nx
        wave_err = max(wave_err, maxval(abs(hx(i, j, :) - 3*cos(3*(i - 1)*dx)*cos(5*(j - 1)*dy)))/3)
        wave_err = max(wave_err, maxval(abs(hy(i, j, :) + 5*sin(3*(i - 1)*dx)*sin(5*(j - 1)*dy)))/5)
      end do
    end do
  end function wave_err

  subroutine report(name, err, code)
    character(len=*), intent(in) :: name
    real(rprec), intent(in) :: err
    integer, intent(in) :: code
    write (*, '(a,a,es10.3)') name, ' max error = ', err
    if (.not. (err <= 1e-12_rprec)) then
      write (*, '(a,a)') 'FAILED: ', name
      error stop code
    end if
  end subroutine report

end program dropin_check
