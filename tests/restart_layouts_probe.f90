! Writes and reads the droplet and temperature restart records the way io.f90:762-771
! (checkpoint_final_v2) and initial_multi.f90:62-74 do: one unformatted sequential record
! per file, arrays in the layouts vel_pcon, vel_sc and vel_sc_pcon, with the reference's
! shapes (ld,ny,nz) volumes, (ld,ny,nz,npcon) bins, sgs_t3(:,:,1), (nx,ny) surfaces and
! (nx,ny,nz) dynamic-Schmidt averages.  Array number q of the record holds
! q*100000 + p*10000 + k*1000 + j*100 + i at (i, j, k, p), so that tests/test_restart.py
! can check the array order and shapes of population_balance_amd/restart.py against a
! Fortran compiler.
! Usage: restart_layouts_probe write|read pcon|sc|scpcon <file>
program restart_layouts_probe
  implicit none
  integer, parameter :: ld = 6, nx = 4, ny = 3, nz = 2, npcon = 2
  real(8), dimension(ld, ny, nz) :: u, v, w, theta, rx, ry, rz, rt, sgs_t3, cs, flm, fmm, fqn, fnn, kct
  real(8), dimension(ld, ny, nz, npcon) :: pcon, rpc
  real(8), dimension(nx, ny) :: psi_m, dep, rdep
  real(8), dimension(nx, ny, nz) :: fkx, fxx, fkx2, fxx2
  character(len=256) :: mode, kind, fname
  integer :: bad
  call get_command_argument(1, mode)
  call get_command_argument(2, kind)
  call get_command_argument(3, fname)
  bad = 0
  if (trim(mode) == 'write') then
    if (trim(kind) == 'pcon') then
      call f3(u, 1); call f3(v, 2); call f3(w, 3); call f4(pcon, 4); call f3(rx, 5); call f3(ry, 6)
      call f3(rz, 7); call f4(rpc, 8); call f3(cs, 9); call f3(flm, 10); call f3(fmm, 11)
      call f3(fqn, 12); call f3(fnn, 13); call f2(dep, 14); call f2(rdep, 15); call f3(kct, 16)
      call fs(fkx, 17); call fs(fxx, 18); call fs(fkx2, 19); call fs(fxx2, 20)
      open(unit=11, file=trim(fname), form='unformatted')
      write (11) u(:, :, 1:nz), v(:, :, 1:nz), w(:, :, 1:nz), pcon(:, :, 1:nz, 1:npcon), &
                 rx(:, :, 1:nz), ry(:, :, 1:nz), rz(:, :, 1:nz), rpc(:, :, 1:nz, 1:npcon), &
                 cs, flm, fmm, fqn, fnn, dep(:, :), rdep, kct, fkx, fxx, fkx2, fxx2
      close(11)
    else if (trim(kind) == 'sc') then
      call f3(u, 1); call f3(v, 2); call f3(w, 3); call f3(theta, 4); call f3(rx, 5); call f3(ry, 6)
      call f3(rz, 7); call f3(rt, 8); call f3(sgs_t3, 9); call f2(psi_m, 10); call f3(cs, 11)
      call f3(flm, 12); call f3(fmm, 13); call f3(fqn, 14); call f3(fnn, 15)
      open(unit=11, file=trim(fname), form='unformatted')
      write (11) u(:, :, 1:nz), v(:, :, 1:nz), w(:, :, 1:nz), theta(:, :, 1:nz), &
                 rx(:, :, 1:nz), ry(:, :, 1:nz), rz(:, :, 1:nz), rt(:, :, 1:nz), sgs_t3(:, :, 1), psi_m, &
                 cs, flm, fmm, fqn, fnn
      close(11)
    else
      call f3(u, 1); call f3(v, 2); call f3(w, 3); call f3(theta, 4); call f4(pcon, 5); call f3(rx, 6)
      call f3(ry, 7); call f3(rz, 8); call f3(rt, 9); call f3(sgs_t3, 10); call f2(psi_m, 11)
      call f4(rpc, 12); call f3(cs, 13); call f3(flm, 14); call f3(fmm, 15); call f3(fqn, 16)
      call f3(fnn, 17); call f2(dep, 18); call f2(rdep, 19); call f3(kct, 20); call fs(fkx, 21)
      call fs(fxx, 22); call fs(fkx2, 23); call fs(fxx2, 24)
      open(unit=11, file=trim(fname), form='unformatted')
      write (11) u(:, :, 1:nz), v(:, :, 1:nz), w(:, :, 1:nz), theta(:, :, 1:nz), &
                 pcon(:, :, 1:nz, 1:npcon), rx(:, :, 1:nz), ry(:, :, 1:nz), rz(:, :, 1:nz), &
                 rt(:, :, 1:nz), sgs_t3(:, :, 1), psi_m, rpc(:, :, 1:nz, 1:npcon), &
                 cs, flm, fmm, fqn, fnn, dep(:, :), rdep, kct, fkx, fxx, fkx2, fxx2
      close(11)
    end if
    print '(a)', 'written'
  else
    open(unit=11, file=trim(fname), form='unformatted', status='old')
    if (trim(kind) == 'pcon') then
      read (11) u(:, :, 1:nz), v(:, :, 1:nz), w(:, :, 1:nz), pcon(:, :, 1:nz, 1:npcon), &
                rx(:, :, 1:nz), ry(:, :, 1:nz), rz(:, :, 1:nz), rpc(:, :, 1:nz, 1:npcon), &
                cs, flm, fmm, fqn, fnn, dep(:, :), rdep, kct, fkx, fxx, fkx2, fxx2
      call c3(u, 1); call c3(v, 2); call c3(w, 3); call c4(pcon, 4); call c3(rx, 5); call c3(ry, 6)
      call c3(rz, 7); call c4(rpc, 8); call c3(cs, 9); call c3(flm, 10); call c3(fmm, 11)
      call c3(fqn, 12); call c3(fnn, 13); call c2(dep, 14); call c2(rdep, 15); call c3(kct, 16)
      call cs3(fkx, 17); call cs3(fxx, 18); call cs3(fkx2, 19); call cs3(fxx2, 20)
    else if (trim(kind) == 'sc') then
      read (11) u(:, :, 1:nz), v(:, :, 1:nz), w(:, :, 1:nz), theta(:, :, 1:nz), &
                rx(:, :, 1:nz), ry(:, :, 1:nz), rz(:, :, 1:nz), rt(:, :, 1:nz), sgs_t3(:, :, 1), psi_m, &
                cs, flm, fmm, fqn, fnn
      call c3(u, 1); call c3(v, 2); call c3(w, 3); call c3(theta, 4); call c3(rx, 5); call c3(ry, 6)
      call c3(rz, 7); call c3(rt, 8); call c3l1(sgs_t3, 9); call c2(psi_m, 10); call c3(cs, 11)
      call c3(flm, 12); call c3(fmm, 13); call c3(fqn, 14); call c3(fnn, 15)
    else
      read (11) u(:, :, 1:nz), v(:, :, 1:nz), w(:, :, 1:nz), theta(:, :, 1:nz), &
                pcon(:, :, 1:nz, 1:npcon), rx(:, :, 1:nz), ry(:, :, 1:nz), rz(:, :, 1:nz), &
                rt(:, :, 1:nz), sgs_t3(:, :, 1), psi_m, rpc(:, :, 1:nz, 1:npcon), &
                cs, flm, fmm, fqn, fnn, dep(:, :), rdep, kct, fkx, fxx, fkx2, fxx2
      call c3(u, 1); call c3(v, 2); call c3(w, 3); call c3(theta, 4); call c4(pcon, 5); call c3(rx, 6)
      call c3(ry, 7); call c3(rz, 8); call c3(rt, 9); call c3l1(sgs_t3, 10); call c2(psi_m, 11)
      call c4(rpc, 12); call c3(cs, 13); call c3(flm, 14); call c3(fmm, 15); call c3(fqn, 16)
      call c3(fnn, 17); call c2(dep, 18); call c2(rdep, 19); call c3(kct, 20); call cs3(fkx, 21)
      call cs3(fxx, 22); call cs3(fkx2, 23); call cs3(fxx2, 24)
    end if
    close(11)
    if (bad == 0) then
      print '(a)', 'read ok'
    else
      print '(a,i0)', 'mismatches ', bad
    end if
  end if
contains
  real(8) function val(q, i, j, k, p)
    integer, intent(in) :: q, i, j, k, p
    val = q * 100000 + p * 10000 + k * 1000 + j * 100 + i
  end function
  subroutine f2(a, q)
    real(8), intent(out) :: a(nx, ny)
    integer, intent(in) :: q
    integer :: i, j
    do j = 1, ny
      do i = 1, nx
        a(i, j) = val(q, i, j, 0, 0)
      end do
    end do
  end subroutine
  subroutine f3(a, q)
    real(8), intent(out) :: a(ld, ny, nz)
    integer, intent(in) :: q
    integer :: i, j, k
    do k = 1, nz
      do j = 1, ny
        do i = 1, ld
          a(i, j, k) = val(q, i, j, k, 0)
        end do
      end do
    end do
  end subroutine
  subroutine fs(a, q)
    real(8), intent(out) :: a(nx, ny, nz)
    integer, intent(in) :: q
    integer :: i, j, k
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          a(i, j, k) = val(q, i, j, k, 0)
        end do
      end do
    end do
  end subroutine
  subroutine f4(a, q)
    real(8), intent(out) :: a(ld, ny, nz, npcon)
    integer, intent(in) :: q
    integer :: i, j, k, p
    do p = 1, npcon
      do k = 1, nz
        do j = 1, ny
          do i = 1, ld
            a(i, j, k, p) = val(q, i, j, k, p)
          end do
        end do
      end do
    end do
  end subroutine
  subroutine c2(a, q)
    real(8), intent(in) :: a(nx, ny)
    integer, intent(in) :: q
    integer :: i, j
    do j = 1, ny
      do i = 1, nx
        if (a(i, j) /= val(q, i, j, 0, 0)) bad = bad + 1
      end do
    end do
  end subroutine
  subroutine c3(a, q)
    real(8), intent(in) :: a(ld, ny, nz)
    integer, intent(in) :: q
    integer :: i, j, k
    do k = 1, nz
      do j = 1, ny
        do i = 1, ld
          if (a(i, j, k) /= val(q, i, j, k, 0)) bad = bad + 1
        end do
      end do
    end do
  end subroutine
  subroutine c3l1(a, q)
    real(8), intent(in) :: a(ld, ny, nz)
    integer, intent(in) :: q
    integer :: i, j
    do j = 1, ny
      do i = 1, ld
        if (a(i, j, 1) /= val(q, i, j, 1, 0)) bad = bad + 1
      end do
    end do
  end subroutine
  subroutine cs3(a, q)
    real(8), intent(in) :: a(nx, ny, nz)
    integer, intent(in) :: q
    integer :: i, j, k
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          if (a(i, j, k) /= val(q, i, j, k, 0)) bad = bad + 1
        end do
      end do
    end do
  end subroutine
  subroutine c4(a, q)
    real(8), intent(in) :: a(ld, ny, nz, npcon)
    integer, intent(in) :: q
    integer :: i, j, k, p
    do p = 1, npcon
      do k = 1, nz
        do j = 1, ny
          do i = 1, ld
            if (a(i, j, k, p) /= val(q, i, j, k, p)) bad = bad + 1
          end do
        end do
      end do
    end do
  end subroutine
end program
