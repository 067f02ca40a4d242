! Writes and reads a restart-shaped record the way io.f90:772-776 and initial.f90:295-297
! do (one unformatted sequential record of eleven REAL(8) (ld,ny,nz) arrays), with
! values q*1000 + k*100 + j*10 + i, so that tests/test_restart.py can check the record
! framing and the array order of population_balance_amd/restart.py against a Fortran
! compiler.  Usage: restart_probe write <file> | restart_probe read <file>
program restart_probe
  implicit none
  integer, parameter :: ld = 6, ny = 4, nz = 3
  real(8), dimension(ld, ny, nz) :: u, v, w, rx, ry, rz, cs, flm, fmm, fqn, fnn
  character(len=256) :: mode, fname
  integer :: bad
  call get_command_argument(1, mode)
  call get_command_argument(2, fname)
  if (trim(mode) == 'write') then
    call fill(u, 1); call fill(v, 2); call fill(w, 3); call fill(rx, 4); call fill(ry, 5)
    call fill(rz, 6); call fill(cs, 7); call fill(flm, 8); call fill(fmm, 9); call fill(fqn, 10)
    call fill(fnn, 11)
    open(unit=11, file=trim(fname), form='unformatted')
    write (11) u(:, :, 1:nz), v(:, :, 1:nz), w(:, :, 1:nz), rx(:, :, 1:nz), ry(:, :, 1:nz), &
               rz(:, :, 1:nz), cs, flm, fmm, fqn, fnn
    close(11)
    print '(a)', 'written'
  else
    open(unit=11, file=trim(fname), form='unformatted', status='old')
    read (11) u(:, :, 1:nz), v(:, :, 1:nz), w(:, :, 1:nz), rx(:, :, 1:nz), ry(:, :, 1:nz), &
              rz(:, :, 1:nz), cs, flm, fmm, fqn, fnn
    close(11)
    bad = 0
    call check(u, 1); call check(v, 2); call check(w, 3); call check(rx, 4); call check(ry, 5)
    call check(rz, 6); call check(cs, 7); call check(flm, 8); call check(fmm, 9)
    call check(fqn, 10); call check(fnn, 11)
    if (bad == 0) then
      print '(a)', 'read ok'
    else
      print '(a,i0)', 'mismatches ', bad
    end if
  end if
contains
  subroutine fill(a, q)
    real(8), intent(out) :: a(ld, ny, nz)
    integer, intent(in) :: q
    integer :: i, j, k
    do k = 1, nz
      do j = 1, ny
        do i = 1, ld
          a(i, j, k) = q * 1000 + k * 100 + j * 10 + i
        end do
      end do
    end do
  end subroutine
  subroutine check(a, q)
    real(8), intent(in) :: a(ld, ny, nz)
    integer, intent(in) :: q
    integer :: i, j, k
    do k = 1, nz
      do j = 1, ny
        do i = 1, ld
          if (a(i, j, k) /= q * 1000 + k * 100 + j * 10 + i) bad = bad + 1
        end do
      end do
    end do
  end subroutine
end program
