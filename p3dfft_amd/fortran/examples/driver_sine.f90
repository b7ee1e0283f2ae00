! driver_sine.f90 -- Fortran usage example of module p3dfft on one GPU.
!
! Same check as the reference's sample/FORTRAN/driver_sine.F90: transform
! sin(2 pi x/nx) sin(2 pi y/ny) sin(2 pi z/nz) forward ('fft'), normalise by
! 1/(nx ny nz), transform back ('tff') and compare with the input; the result
! is correct when the maximum difference is below prec * nx*ny*nz / 4.
! Usage: driver_sine_f.x nx ny nz niter
program driver_sine
  use p3dfft
  implicit none
  integer :: nx, ny, nz, niter, it, x, y, z, dims(2), memsize(3)
  integer :: istart(3), iend(3), isize(3), fstart(3), fend(3), fsize(3)
  real(p3dfft_type), allocatable :: a(:,:,:), c(:,:,:)
  complex(p3dfft_type), allocatable :: b(:,:,:)
  real(r8) :: twopi, cdiff, prec, factor, t(12)
  character(len=32) :: arg

  nx = 64; ny = 64; nz = 64; niter = 2
  if (command_argument_count() >= 4) then
    call get_command_argument(1, arg); read(arg, *) nx
    call get_command_argument(2, arg); read(arg, *) ny
    call get_command_argument(3, arg); read(arg, *) nz
    call get_command_argument(4, arg); read(arg, *) niter
  end if
  twopi = 8.0_r8 * atan(1.0_r8)
  dims = (/1, 1/)
  call p3dfft_setup(dims, nx, ny, nz, 0, nx, ny, nz, .true., memsize)
  call p3dfft_get_dims(istart, iend, isize, 1)
  call p3dfft_get_dims(fstart, fend, fsize, 2)
  allocate(a(isize(1), isize(2), isize(3)), c(isize(1), isize(2), isize(3)))
  allocate(b(fsize(1), fsize(2), fsize(3)))
  do z = 1, isize(3)
    do y = 1, isize(2)
      do x = 1, isize(1)
        a(x, y, z) = sin(twopi * (x - 1) / nx) * sin(twopi * (y - 1) / ny) * sin(twopi * (z - 1) / nz)
      end do
    end do
  end do
  call p3dfft_ftran_r2c(a, b, 'fft')
  call set_timers()
  factor = 1.0_r8 / (real(nx, r8) * real(ny, r8) * real(nz, r8))
  do it = 1, niter
    call p3dfft_ftran_r2c(a, b, 'fft')
    b = b * factor
    call p3dfft_btran_c2r(b, c, 'tff')
  end do
  cdiff = maxval(abs(c - a))
  call get_timers(t)
  call p3dfft_clean()
  if (p3dfft_type == 8) then
    prec = 1e-14_r8
  else
    prec = 1e-5_r8
  end if
  print *, 'max diff =', cdiff
  print *, 'timers 5,7,8,9,10,12 (s):', t(5), t(7), t(8), t(9), t(10), t(12)
  if (cdiff > prec * real(nx, r8) * real(ny, r8) * real(nz, r8) * 0.25_r8) then
    print *, 'Results are incorrect'
  else
    print *, 'Results are correct'
  end if
end program driver_sine
