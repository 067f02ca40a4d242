! drop_in_demo.f90 -- a Fortran host calling the HIP implementation through the
! ISO_C_BINDING module, with arrays declared exactly as the reference declares them
! (real(rprec) f(ld,ny,nz), column-major).  Without arguments it only checks the
! structure layout; with "--run" it creates a context on the GPU and checks filt_da
! (filt_da.f90 argument order), ddz_uv and the per-level spectrum of io.f90:1643-1679 on
! closed-form fields.
program drop_in_demo
  use iso_c_binding
  use lesgo_amd
  implicit none
  integer, parameter :: rprec = c_double
  type(lp_config), target :: cfg
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  integer :: jx, jy, jz, nx, ny, nz, ld
  real(rprec), allocatable :: f(:,:,:), dfdx(:,:,:), dfdy(:,:,:), dfdz(:,:,:)
  real(rprec), allocatable, target :: fu(:,:,:)
  real(rprec), allocatable :: spec(:,:)
  real(rprec) :: x, y, err, pi
  character(len=16) :: arg

  call lp_config_default(cfg)
  rc = lp_grid_init(cfg%grid, 32_c_int32_t, 32_c_int32_t, 9_c_int32_t, &
                    8.0_rprec*atan(1.0_rprec), 8.0_rprec*atan(1.0_rprec), 1.0_rprec)
  write (*, '(a,i0)') 'sizeof(lp_config) = ', c_sizeof(cfg)
  write (*, '(a,3i6,a,i0)') 'grid ', cfg%grid%nx, cfg%grid%ny, cfg%grid%nz, '  ld = ', cfg%grid%ld
  arg = ' '
  if (command_argument_count() >= 1) call get_command_argument(1, arg)
  if (trim(arg) /= '--run') stop 0

  rc = lp_create(cfg, ctx)
  if (rc /= LP_OK) then
    write (*, *) 'lp_create failed: ', trim(lp_last_error_f())
    stop 1
  end if
  nx = cfg%grid%nx; ny = cfg%grid%ny; nz = cfg%grid%nz; ld = cfg%grid%ld
  allocate (f(ld, ny, nz), dfdx(ld, ny, nz), dfdy(ld, ny, nz), dfdz(ld, ny, nz))
  pi = 4.0_rprec*atan(1.0_rprec)
  f = 0; dfdx = 0; dfdy = 0; dfdz = 0
  do jz = 1, nz
    do jy = 1, ny
      do jx = 1, nx
        x = (jx - 1)*cfg%grid%dx
        y = (jy - 1)*cfg%grid%dy
        f(jx, jy, jz) = sin(3*x)*cos(5*y) + jz
      end do
    end do
  end do
  rc = lp_filt_da(ctx, f, dfdx, dfdy, int(nz, c_int32_t))
  if (rc /= LP_OK) stop 2
  err = 0
  do jz = 1, nz
    do jy = 1, ny
      do jx = 1, nx
        x = (jx - 1)*cfg%grid%dx
        y = (jy - 1)*cfg%grid%dy
        err = max(err, abs(dfdx(jx, jy, jz) - 3*cos(3*x)*cos(5*y)))
        err = max(err, abs(dfdy(jx, jy, jz) + 5*sin(3*x)*sin(5*y)))
      end do
    end do
  end do
  write (*, '(a,es10.3)') 'filt_da max error = ', err
  if (err > 1e-12) stop 8
  rc = lp_ddz_uv(ctx, dfdz, f)
  if (rc /= LP_OK) stop 3
  err = maxval(abs(dfdz(1:nx, :, 2:nz-1) - 1.0_rprec/cfg%grid%dz))
  write (*, '(a,es10.3)') 'ddz_uv max error = ', err*cfg%grid%dz
  if (err*cfg%grid%dz > 1e-12) stop 4
  ! spectrum of u = sin(3x) cos(5y) + jz: the row mean jz gives E(k=0) = jz**2/2, the
  ! sin(3x) wave of amplitude cos(5y) gives E(k=3) = <cos(5y)**2>/4 = 1/8, nothing else
  allocate (fu(ld, ny, nz), spec(nx/2, nz-1))
  fu = 0
  do jz = 1, nz
    do jy = 1, ny
      do jx = 1, nx
        x = (jx - 1)*cfg%grid%dx
        y = (jy - 1)*cfg%grid%dy
        fu(jx, jy, jz) = sin(3*x)*cos(5*y) + jz
      end do
    end do
  end do
  rc = lp_upload(ctx, LP_U, c_loc(fu), int(8, c_size_t)*ld*ny*nz)
  if (rc /= LP_OK) stop 5
  rc = lp_spectrum(ctx, LP_U, 0_c_int32_t, spec)
  if (rc /= LP_OK) stop 6
  err = 0
  do jz = 1, nz - 1
    do jx = 1, nx/2
      x = 0
      if (jx == 1) x = 0.5_rprec*jz*jz
      if (jx == 4) x = 0.125_rprec
      err = max(err, abs(spec(jx, jz) - x))
    end do
  end do
  write (*, '(a,es10.3)') 'spectrum max error = ', err
  rc = lp_destroy(ctx)
  if (err > 1e-12) stop 7
  write (*, '(a)') 'drop-in demo ok'
end program drop_in_demo
