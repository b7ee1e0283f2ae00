! wrap.f90 -- external (non-module) transform entry points for in-place use
! from Fortran: the same input and output array can be passed, as with the
! reference's build/wrap.F90:82-150.
subroutine ftran_r2c(IN, OUT, op)
  use p3dfft
  implicit none
  real(p3dfft_type), target :: IN(*)
  complex(p3dfft_type), target :: OUT(*)
  character(len=3) :: op
  call p3dfft_ftran_r2c(IN, OUT, op)
end subroutine

subroutine btran_c2r(IN, OUT, op)
  use p3dfft
  implicit none
  complex(p3dfft_type), target :: IN(*)
  real(p3dfft_type), target :: OUT(*)
  character(len=3) :: op
  call p3dfft_btran_c2r(IN, OUT, op)
end subroutine

subroutine ftran_r2c_many(IN, dim_in, OUT, dim_out, nv, op)
  use p3dfft
  implicit none
  integer :: dim_in, dim_out, nv
  real(p3dfft_type), target :: IN(*)
  complex(p3dfft_type), target :: OUT(*)
  character(len=3) :: op
  call p3dfft_ftran_r2c_many(IN, dim_in, OUT, dim_out, nv, op)
end subroutine

subroutine btran_c2r_many(IN, dim_in, OUT, dim_out, nv, op)
  use p3dfft
  implicit none
  integer :: dim_in, dim_out, nv
  complex(p3dfft_type), target :: IN(*)
  real(p3dfft_type), target :: OUT(*)
  character(len=3) :: op
  call p3dfft_btran_c2r_many(IN, dim_in, OUT, dim_out, nv, op)
end subroutine
