! driver_utils.f90 -- the module's utility procedures on one GPU: the MPI
! information, the real-data pencil transposes (copies of the whole grid on
! one rank) and print_buf of a spectrum with two non-zero modes.  Prints
! "Utilities are correct" when every check passes.
program driver_utils
  use p3dfft
  implicit none
  integer, parameter :: nx = 16, ny = 12, nz = 8
  integer :: dims(2), st(3), en(3), sz(3), ds(3), de(3), dz(3), tid, ntask, comm, i, j, k, bad
  real(p3dfft_type), allocatable :: a(:, :, :), b(:, :, :), r1(:), r2(:)
  complex(p3dfft_type), allocatable :: c(:, :, :)
  real(r8) :: t
  dims = (/1, 1/)
  call p3dfft_setup(dims, nx, ny, nz, 4321)
  call p3dfft_get_dims(st, en, sz, 1)
  allocate (a(sz(1), sz(2), sz(3)), b(sz(1), sz(2), sz(3)), r1(1), r2(1))
  do k = 1, sz(3)
    do j = 1, sz(2)
      do i = 1, sz(1)
        a(i, j, k) = real(i + 100 * j + 10000 * k, p3dfft_type)
      end do
    end do
  end do
  bad = 0
  call p3dfft_get_mpi_info(tid, ntask, comm)
  if (tid /= 0 .or. ntask /= 1 .or. comm /= 4321) bad = bad + 1
  t = 0.0_r8
  b = 0
  call rtran_x2y(a, b, r1, r2, ds, de, dz, t)
  if (any(b /= a) .or. any(dz /= (/nx, ny, nz/)) .or. any(ds /= 1) .or. any(de /= (/nx, ny, nz/))) bad = bad + 1
  b = 0
  call rtran_y2x(a, b, r1, r2, ds, de, dz, t)
  if (any(b /= a)) bad = bad + 1
  b = 0
  call rtran_x2z(a, b, r1, r2, ds, de, dz, t)
  if (any(b /= a)) bad = bad + 1
  b = 0
  call rtran_z2x(a, b, r1, r2, ds, de, dz, t)
  if (any(b /= a) .or. t /= 0.0_r8) bad = bad + 1
  call p3dfft_get_dims(st, en, sz, 2)
  allocate (c(sz(1), sz(2), sz(3)))
  c = (0.0, 0.0)
  c(2, 3, 4) = (1.5, -2.0)
  c(5, 1, 1) = (0.0, 3.0)
  call print_buf(c, sz(1), sz(2), sz(3))
  call p3dfft_clean()
  if (bad == 0) then
    print *, 'Utilities are correct'
  else
    print *, 'Utilities failed:', bad
  end if
end program
