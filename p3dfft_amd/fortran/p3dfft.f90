! p3dfft.f90 -- Fortran module `p3dfft` of libp3dfft_amd.
!
! Same module name, kinds and procedure names as the reference's module
! (build/module.F90:81-186) for the ftran_r2c / btran_c2r path, with the
! optional arguments of p3dfft_setup (build/setup.F90:107-120).  Every call
! forwards through iso_c_binding to the C-ABI of libp3dfft_amd.so (or, built
! with -DSINGLE_PREC, libp3dfft_amd_sp.so).  Arrays are host arrays; the
! library copies them to and from HBM inside each call.
module p3dfft
  use, intrinsic :: iso_c_binding
  implicit none
  private

#ifndef SINGLE_PREC
  integer, parameter, public :: p3dfft_type = kind(1.0d0)
#else
  integer, parameter, public :: p3dfft_type = kind(1.0)
#endif
  integer, parameter, public :: r8 = kind(1.0d0)
  integer, parameter, public :: i8 = selected_int_kind(18)
  ! stage timers (build/module.F90:106), refreshed after every transform call
  real(r8), save, public :: timers(12) = 0.0_r8
  ! MPI datatypes of p3dfft_type reals and complexes (build/module.F90:95-104).
  ! This library does not link MPI: they hold -1 until a caller that uses MPI
  ! sets them with p3dfft_set_mpi_types (e.g. MPI_DOUBLE_PRECISION and
  ! MPI_DOUBLE_COMPLEX), after which they can be used as in the reference.
  integer, save, public :: p3dfft_mpireal = -1, p3dfft_mpicomplex = -1

  ! processor-grid geometry of the setup (build/module.F90:172-176): every
  ! rank's get_dims extents proc_dims(conf, field, rank) (fields: start x,y,z,
  ! end x,y,z, size x,y,z), the rank <-> (ipid, jpid) maps and the result
  ! array of get_proc_parts
  integer, save, dimension(:), allocatable, public :: proc_id2coords
  integer, save, dimension(:, :), allocatable, public :: proc_coords2id
  integer, save, dimension(:, :, :), allocatable, public :: proc_dims
  integer, save, dimension(:, :), allocatable, public :: proc_parts

  public :: p3dfft_setup, p3dfft_get_dims, p3dfft_ftran_r2c, p3dfft_btran_c2r, &
            p3dfft_cheby, p3dfft_ftran_r2c_many, p3dfft_btran_c2r_many, p3dfft_cheby_many, &
            p3dfft_ftran_r2c_1d, p3dfft_clean, get_timers, set_timers, p3dfft_set_mpi_types, &
            p3dfft_get_mpi_info, print_buf, print_buf_real, rtran_x2y, rtran_y2x, rtran_x2z, rtran_z2x, &
            get_proc_parts

  ! the communicator handed to p3dfft_setup (one GPU is task 0 of 1)
  integer, save :: setup_comm = -1
  logical, save :: setup_done = .false.

  interface
    subroutine c_p3dfft_setup(dims, nx, ny, nz, comm, nxc, nyc, nzc, ow, memsize) &
        bind(C, name='p3dfft_setup')
      import :: c_int
      integer(c_int) :: dims(2), nx, ny, nz, comm, nxc, nyc, nzc, ow, memsize(3)
    end subroutine
    subroutine c_p3dfft_get_dims(st, en, sz, conf) bind(C, name='p3dfft_get_dims')
      import :: c_int
      integer(c_int) :: st(3), en(3), sz(3), conf
    end subroutine
    subroutine c_p3dfft_ftran_r2c(a, b, op) bind(C, name='p3dfft_ftran_r2c')
      import :: c_ptr, c_char
      type(c_ptr), value :: a, b
      character(kind=c_char) :: op(*)
    end subroutine
    subroutine c_p3dfft_btran_c2r(a, b, op) bind(C, name='p3dfft_btran_c2r')
      import :: c_ptr, c_char
      type(c_ptr), value :: a, b
      character(kind=c_char) :: op(*)
    end subroutine
    subroutine c_p3dfft_cheby(a, b, lz) bind(C, name='p3dfft_cheby')
      import :: c_ptr
      type(c_ptr), value :: a, b, lz
    end subroutine
    subroutine c_p3dfft_ftran_r2c_many(a, dim_in, b, dim_out, nv, op) bind(C, name='p3dfft_ftran_r2c_many')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: a, b
      integer(c_int) :: dim_in, dim_out, nv
      character(kind=c_char) :: op(*)
    end subroutine
    subroutine c_p3dfft_btran_c2r_many(a, dim_in, b, dim_out, nv, op) bind(C, name='p3dfft_btran_c2r_many')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: a, b
      integer(c_int) :: dim_in, dim_out, nv
      character(kind=c_char) :: op(*)
    end subroutine
    subroutine c_p3dfft_cheby_many(a, dim_in, b, dim_out, nv, lz) bind(C, name='p3dfft_cheby_many')
      import :: c_ptr, c_int
      type(c_ptr), value :: a, b, lz
      integer(c_int) :: dim_in, dim_out, nv
    end subroutine
    subroutine c_p3dfft_ftran_r2c_1d(a, b) bind(C, name='p3dfft_ftran_r2c_1d')
      import :: c_ptr
      type(c_ptr), value :: a, b
    end subroutine
    subroutine c_p3dfft_clean() bind(C, name='p3dfft_clean')
    end subroutine
    subroutine c_get_timers(t) bind(C, name='get_timers')
      import :: c_double
      real(c_double) :: t(12)
    end subroutine
    subroutine c_set_timers() bind(C, name='set_timers')
    end subroutine
    function c_proc_dims(d, c) bind(C, name='p3dfft_amd_proc_dims') result(np)
      import :: c_ptr, c_int
      type(c_ptr), value :: d, c
      integer(c_int) :: np
    end function
    function c_get_proc_parts(bx, by, bz, sx, sy, sz, conf, parts, ierr) &
        bind(C, name='p3dfft_amd_get_proc_parts') result(n)
      import :: c_ptr, c_int
      integer(c_int), value :: bx, by, bz, sx, sy, sz, conf
      type(c_ptr), value :: parts, ierr
      integer(c_int) :: n
    end function
  end interface

contains

  ! build/setup.F90:107-120: cut sizes default to the grid, overwrite to .true.
  subroutine p3dfft_setup(dims, nx, ny, nz, mpi_comm_in, nxcut, nycut, nzcut, overwrite, memsize)
    integer, intent(in) :: dims(2), nx, ny, nz, mpi_comm_in
    integer, intent(in), optional :: nxcut, nycut, nzcut
    logical, intent(in), optional :: overwrite
    integer, intent(out), optional :: memsize(3)
    integer(c_int) :: d(2), cnx, cny, cnz, comm, cx, cy, cz, ow, ms(3)
    d = dims
    cnx = nx; cny = ny; cnz = nz; comm = mpi_comm_in
    cx = nx; cy = ny; cz = nz; ow = 1
    if (present(nxcut)) cx = nxcut
    if (present(nycut)) cy = nycut
    if (present(nzcut)) cz = nzcut
    if (present(overwrite)) then
      if (.not. overwrite) ow = 0
    end if
    call c_p3dfft_setup(d, cnx, cny, cnz, comm, cx, cy, cz, ow, ms)
    if (present(memsize)) memsize = ms
    timers = 0.0_r8
    setup_comm = mpi_comm_in
    setup_done = .true.
    call fill_proc_arrays(dims(1), dims(2))
  end subroutine

  ! build/setup.F90:224-229, 551-574 (the MPI_Allgather of every rank's
  ! coordinates and extents), from the library's own record of the splits
  subroutine fill_proc_arrays(iproc, jproc)
    integer, intent(in) :: iproc, jproc
    integer(c_int), target, allocatable :: d(:), c(:)
    integer :: np, p
    call free_proc_arrays()
    np = c_proc_dims(c_null_ptr, c_null_ptr)
    if (np < 1 .or. np /= iproc * jproc) return
    allocate (d(18 * np), c(2 * np))
    np = c_proc_dims(c_loc(d(1)), c_loc(c(1)))
    allocate (proc_dims(2, 9, 0:np - 1), proc_id2coords(0:2 * np - 1), &
              proc_coords2id(0:iproc - 1, 0:jproc - 1), proc_parts(np, 7))
    proc_dims = reshape(d, (/2, 9, np/))
    proc_id2coords = c
    do p = 0, np - 1
      proc_coords2id(c(2 * p + 1), c(2 * p + 2)) = p
    end do
    proc_parts = -1
  end subroutine

  subroutine free_proc_arrays()
    if (allocated(proc_dims)) deallocate (proc_dims)
    if (allocated(proc_id2coords)) deallocate (proc_id2coords)
    if (allocated(proc_coords2id)) deallocate (proc_coords2id)
    if (allocated(proc_parts)) deallocate (proc_parts)
  end subroutine

  ! build/module.F90:888-1054: proc_parts(q, :) = (rank, base x, y, z,
  ! size x, y, z) of the q-th piece of the box held by one rank; unused rows
  ! -1.  ierr: 0, 1 (conf not 1 or 2), -1 (box not inside the grid).
  subroutine get_proc_parts(base_x, base_y, base_z, size_x, size_y, size_z, conf, ierr)
    integer, intent(in) :: base_x, base_y, base_z, size_x, size_y, size_z, conf
    integer, intent(out) :: ierr
    integer(c_int), target, allocatable :: a(:)
    integer(c_int), target :: e
    integer(c_int) :: n
    integer :: np
    ierr = -1
    if (.not. allocated(proc_parts)) return
    np = size(proc_parts, 1)
    allocate (a(7 * np))
    e = 0
    n = c_get_proc_parts(base_x, base_y, base_z, size_x, size_y, size_z, conf, c_loc(a(1)), c_loc(e))
    proc_parts = reshape(a, (/np, 7/))
    ierr = e
  end subroutine

  ! build/module.F90:225-273
  subroutine p3dfft_get_dims(istart, iend, isize, conf)
    integer, intent(out) :: istart(3), iend(3), isize(3)
    integer, intent(in) :: conf
    integer(c_int) :: s(3), e(3), z(3), c
    c = conf
    call c_p3dfft_get_dims(s, e, z, c)
    istart = s; iend = e; isize = z
  end subroutine

  ! build/ftran.F90:489 (arrays of any rank by sequence association)
  subroutine p3dfft_ftran_r2c(XgYZ, XYZg, op)
    real(p3dfft_type), target :: XgYZ(*)
    complex(p3dfft_type), target :: XYZg(*)
    character(len=*), intent(in) :: op
    call c_p3dfft_ftran_r2c(c_loc(XgYZ(1)), c_loc(XYZg(1)), op(1:3)//c_null_char)
    call c_get_timers(timers)
  end subroutine

  ! build/btran.F90:396
  subroutine p3dfft_btran_c2r(XYZg, XgYZ, op)
    complex(p3dfft_type), target :: XYZg(*)
    real(p3dfft_type), target :: XgYZ(*)
    character(len=*), intent(in) :: op
    call c_p3dfft_btran_c2r(c_loc(XYZg(1)), c_loc(XgYZ(1)), op(1:3)//c_null_char)
    call c_get_timers(timers)
  end subroutine

  ! build/ftran.F90:383
  subroutine p3dfft_cheby(in, out, Lz)
    real(p3dfft_type), target :: in(*)
    complex(p3dfft_type), target :: out(*)
    real(p3dfft_type), intent(in) :: Lz
    real(p3dfft_type), target :: lzc
    lzc = Lz
    call c_p3dfft_cheby(c_loc(in(1)), c_loc(out(1)), c_loc(lzc))
    call c_get_timers(timers)
  end subroutine

  ! build/ftran.F90:104 (variable j of XgYZ at j*dim_in reals, of XYZg at
  ! j*dim_out complex elements)
  subroutine p3dfft_ftran_r2c_many(XgYZ, dim_in, XYZg, dim_out, nv, op)
    integer, intent(in) :: dim_in, dim_out, nv
    real(p3dfft_type), target :: XgYZ(*)
    complex(p3dfft_type), target :: XYZg(*)
    character(len=*), intent(in) :: op
    integer(c_int) :: di, do_, n
    di = dim_in; do_ = dim_out; n = nv
    call c_p3dfft_ftran_r2c_many(c_loc(XgYZ(1)), di, c_loc(XYZg(1)), do_, n, op(1:3)//c_null_char)
    call c_get_timers(timers)
  end subroutine

  ! build/btran.F90:104
  subroutine p3dfft_btran_c2r_many(XYZg, dim_in, XgYZ, dim_out, nv, op)
    integer, intent(in) :: dim_in, dim_out, nv
    complex(p3dfft_type), target :: XYZg(*)
    real(p3dfft_type), target :: XgYZ(*)
    character(len=*), intent(in) :: op
    integer(c_int) :: di, do_, n
    di = dim_in; do_ = dim_out; n = nv
    call c_p3dfft_btran_c2r_many(c_loc(XYZg(1)), di, c_loc(XgYZ(1)), do_, n, op(1:3)//c_null_char)
    call c_get_timers(timers)
  end subroutine

  ! build/ftran.F90:339
  subroutine p3dfft_cheby_many(in, dim_in, out, dim_out, nv, Lz)
    integer, intent(in) :: dim_in, dim_out, nv
    real(p3dfft_type), target :: in(*)
    complex(p3dfft_type), target :: out(*)
    real(p3dfft_type), intent(in) :: Lz
    real(p3dfft_type), target :: lzc
    integer(c_int) :: di, do_, n
    di = dim_in; do_ = dim_out; n = nv; lzc = Lz
    call c_p3dfft_cheby_many(c_loc(in(1)), di, c_loc(out(1)), do_, n, c_loc(lzc))
    call c_get_timers(timers)
  end subroutine

  ! build/ftran.F90:787-814: X r2c only; cXgYZ holds (nx/2+1) complex per row
  subroutine p3dfft_ftran_r2c_1d(rXgYZ, cXgYZ)
    real(p3dfft_type), target :: rXgYZ(*), cXgYZ(*)
    call c_p3dfft_ftran_r2c_1d(c_loc(rXgYZ(1)), c_loc(cXgYZ(1)))
  end subroutine

  subroutine p3dfft_set_mpi_types(real_type, complex_type)
    integer, intent(in) :: real_type, complex_type
    p3dfft_mpireal = real_type
    p3dfft_mpicomplex = complex_type
  end subroutine

  ! build/module.F90:309
  subroutine p3dfft_clean()
    call c_p3dfft_clean()
    call free_proc_arrays()
    setup_done = .false.
  end subroutine

  ! build/module.F90:280-297 on one GPU: task 0 of 1, the setup's communicator
  subroutine p3dfft_get_mpi_info(mpi_taskid, mpi_tasks, mpi_comm)
    integer, intent(out) :: mpi_taskid, mpi_tasks, mpi_comm
    if (.not. setup_done) then
      print *, 'P3DFFT error: call setup before other routines'
      return
    end if
    mpi_taskid = 0
    mpi_tasks = 1
    mpi_comm = setup_comm
  end subroutine

  ! build/module.F90:747-781: every entry above 5e-7 in magnitude, with its
  ! indices, as task 0
  subroutine print_buf(A, lx, ly, lz)
    integer, intent(in) :: lx, ly, lz
    complex(p3dfft_type), intent(in) :: A(lx, ly, lz)
    integer :: i, j, k
    do k = 1, lz
      do j = 1, ly
        do i = 1, lx
          if (abs(A(i, j, k)) > 0.0000005) print *, 0, ': (', i, j, k, ') =', A(i, j, k)
        end do
      end do
    end do
  end subroutine

  subroutine print_buf_real(A, lx, ly, lz)
    integer, intent(in) :: lx, ly, lz
    real(p3dfft_type), intent(in) :: A(lx, ly, lz)
    integer :: i, j, k
    do k = 1, lz
      do j = 1, ly
        do i = 1, lx
          if (abs(A(i, j, k)) > 0.0000005) print *, 0, ': (', i, j, k, ') =', A(i, j, k)
        end do
      end do
    end do
  end subroutine

  ! build/module.F90:1061-1361, the real-data pencil transposes.  On one GPU
  ! the X, Y and Z pencils of real data are all the whole (nx, ny, nz) grid,
  ! so each is a copy with the grid's bounds; the exchange buffers and the
  ! exchange time t are untouched.
  subroutine rtran_x2y(source, dest, rbuf1, rbuf2, dstart, dend, dsize, t)
    real(p3dfft_type), intent(in) :: source(*)
    real(p3dfft_type), intent(out) :: dest(*)
    real(p3dfft_type), intent(inout) :: rbuf1(*), rbuf2(*)
    integer, intent(out) :: dstart(3), dend(3), dsize(3)
    real(r8) :: t
    call rtran_whole(source, dest, dstart, dend, dsize)
  end subroutine

  subroutine rtran_y2x(source, dest, rbuf1, rbuf2, dstart, dend, dsize, t)
    real(p3dfft_type), intent(in) :: source(*)
    real(p3dfft_type), intent(out) :: dest(*)
    real(p3dfft_type), intent(inout) :: rbuf1(*), rbuf2(*)
    integer, intent(out) :: dstart(3), dend(3), dsize(3)
    real(r8) :: t
    call rtran_whole(source, dest, dstart, dend, dsize)
  end subroutine

  subroutine rtran_x2z(source, dest, rbuf1, rbuf2, dstart, dend, dsize, t)
    real(p3dfft_type), intent(in) :: source(*)
    real(p3dfft_type), intent(out) :: dest(*)
    real(p3dfft_type), intent(inout) :: rbuf1(*), rbuf2(*)
    integer, intent(out) :: dstart(3), dend(3), dsize(3)
    real(r8) :: t
    call rtran_whole(source, dest, dstart, dend, dsize)
  end subroutine

  subroutine rtran_z2x(source, dest, rbuf1, rbuf2, dstart, dend, dsize, t)
    real(p3dfft_type), intent(in) :: source(*)
    real(p3dfft_type), intent(out) :: dest(*)
    real(p3dfft_type), intent(inout) :: rbuf1(*), rbuf2(*)
    integer, intent(out) :: dstart(3), dend(3), dsize(3)
    real(r8) :: t
    call rtran_whole(source, dest, dstart, dend, dsize)
  end subroutine

  subroutine rtran_whole(source, dest, dstart, dend, dsize)
    real(p3dfft_type), intent(in) :: source(*)
    real(p3dfft_type), intent(out) :: dest(*)
    integer, intent(out) :: dstart(3), dend(3), dsize(3)
    integer(i8) :: n
    call p3dfft_get_dims(dstart, dend, dsize, 1)
    n = int(dsize(1), i8) * dsize(2) * dsize(3)
    dest(1:n) = source(1:n)
  end subroutine

  ! build/module.F90:730-744
  subroutine get_timers(timer)
    real(r8), intent(out) :: timer(12)
    call c_get_timers(timer)
    timers = timer
  end subroutine

  subroutine set_timers()
    call c_set_timers()
    timers = 0.0_r8
  end subroutine

end module p3dfft
