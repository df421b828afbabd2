!> Kinds of infrastructure/source/utilities/constants_mod.F90:55-61 used by the solver API
!> (r_def = real64, r_solver = real32 as in the default build, i_def = int32).
module constants_mod
  use, intrinsic :: iso_fortran_env, only : real32, real64, int32
  implicit none
  private
  integer, parameter, public :: i_def = int32
  integer, parameter, public :: r_def = real64
  integer, parameter, public :: r_double = real64
  integer, parameter, public :: r_single = real32
  integer, parameter, public :: r_solver = real32
  integer, parameter, public :: l_def = kind(.true.)
  real(r_def), parameter, public :: EPS = epsilon(1.0_r_def)
end module constants_mod
