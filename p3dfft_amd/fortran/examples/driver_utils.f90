! driver_utils.f90 -- the module's utility procedures on one GPU: the MPI
! information, the real-data pencil transposes (copies of the whole grid on
! one rank), the processor-grid arrays and get_proc_parts of the 1x1 grid,
! and print_buf of a spectrum with two non-zero modes.  Prints "Utilities
! are correct" when every check passes.
program driver_utils
  use p3dfft
  implicit none
  integer, parameter :: nx = 16, ny = 12, nz = 8
  integer :: dims(2), st(3), en(3), sz(3), ds(3), de(3), dz(3), tid, ntask, comm, i, j, k, bad, ierr
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
  ! one rank: its extents in both configurations, id 0 at coordinates (0, 0)
  if (size(proc_dims, 3) /= 1 .or. any(proc_dims(1, 7:9, 0) /= (/nx, ny, nz/))) bad = bad + 1
  if (any(proc_dims(2, 1:3, 0) /= 1) .or. any(proc_dims(2, 4:6, 0) /= (/nx / 2 + 1, ny, nz/))) bad = bad + 1
  if (any(proc_id2coords /= 0) .or. proc_coords2id(0, 0) /= 0) bad = bad + 1
  ! a box inside the grid is one piece on rank 0; one that leaves it is refused
  call get_proc_parts(3, 2, 5, 10, 7, 4, 1, ierr)
  if (ierr /= 0 .or. any(proc_parts(1, :) /= (/0, 3, 2, 5, 10, 7, 4/))) bad = bad + 1
  call get_proc_parts(2, 1, 1, 8, ny, 3, 2, ierr)
  if (ierr /= 0 .or. any(proc_parts(1, :) /= (/0, 2, 1, 1, 8, ny, 3/))) bad = bad + 1
  call get_proc_parts(1, ny, 1, nx, 2, nz, 1, ierr)
  if (ierr /= -1 .or. any(proc_parts /= -1)) bad = bad + 1
  call get_proc_parts(1, 1, 1, 1, 1, 1, 3, ierr)
  if (ierr /= 1) bad = bad + 1
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
