! driver_sine_inplace.f90 -- in-place use of the transforms from Fortran
! through the external ftran_r2c / btran_c2r (wrap.f90, the reference's
! build/wrap.F90:82-150), as the reference's
! sample/FORTRAN/driver_sine_inplace.F90 does: one real array, sized by the
! memsize that p3dfft_setup returns, is passed as both the input and the
! output of the forward transform ('fft') and then of the backward one
! ('tff').  Checks the four non-zero modes of sin(2 pi x/nx) sin(2 pi y/ny)
! sin(2 pi z/nz) in the stored half spectrum (each of magnitude nx ny nz / 8)
! and the round trip against the input (prec * nx ny nz / 4).  Prints
! "Results are correct" when both hold.
! Usage: driver_sine_inplace_f.x nx ny nz niter
program driver_sine_inplace
  use p3dfft
  implicit none
  external :: ftran_r2c, btran_c2r
  integer :: nx, ny, nz, niter, it, x, y, z, dims(2), memsize(3), nbig, bad
  integer :: istart(3), iend(3), isize(3)
  integer(i8) :: ntot, nc, i
  real(p3dfft_type), allocatable :: B(:), A(:)
  real(r8) :: twopi, cdiff, prec, factor, mag, big
  character(len=32) :: arg

  nx = 32; ny = 24; nz = 20; niter = 2
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
  ! the in-place array: memsize(1) x memsize(2) x memsize(3) reals, which
  ! holds the real field and the (nx/2+1) x ny x nz complex spectrum
  allocate (B(int(memsize(1), i8) * memsize(2) * memsize(3)))
  ntot = int(isize(1), i8) * isize(2) * isize(3)
  allocate (A(ntot))
  do z = 1, isize(3)
    do y = 1, isize(2)
      do x = 1, isize(1)
        A(x + isize(1) * ((y - 1) + isize(2) * (z - 1))) = &
          sin(twopi * (x - 1) / nx) * sin(twopi * (y - 1) / ny) * sin(twopi * (z - 1) / nz)
      end do
    end do
  end do
  factor = 1.0_r8 / (real(nx, r8) * real(ny, r8) * real(nz, r8))
  nc = int(nx / 2 + 1, i8) * ny * nz
  bad = 0
  B(1:ntot) = A
  do it = 1, niter
    call ftran_r2c(B, B, 'fft')
    if (it == 1) then
      ! the spectrum sits in B as (re, im) pairs
      nbig = 0
      big = real(nx, r8) * real(ny, r8) * real(nz, r8) / 8.0_r8
      do i = 1, nc
        mag = sqrt(real(B(2 * i - 1), r8) ** 2 + real(B(2 * i), r8) ** 2)
        if (mag > 0.5_r8 * big) then
          nbig = nbig + 1
          if (abs(mag - big) > 1e-3_r8 * big) bad = bad + 1
        end if
      end do
      if (nbig /= 4) bad = bad + 1
      print *, 'modes above half the expected magnitude:', nbig
    end if
    B(1:2 * nc) = B(1:2 * nc) * factor
    call btran_c2r(B, B, 'tff')
  end do
  cdiff = maxval(abs(B(1:ntot) - A))
  call p3dfft_clean()
  if (p3dfft_type == 8) then
    prec = 1e-14_r8
  else
    prec = 1e-5_r8
  end if
  print *, 'max diff =', cdiff
  if (bad == 0 .and. cdiff <= prec * real(nx, r8) * real(ny, r8) * real(nz, r8) * 0.25_r8) then
    print *, 'Results are correct'
  else
    print *, 'Results are incorrect', bad
  end if
end program driver_sine_inplace
