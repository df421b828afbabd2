!> Interface of abstract_iterative_solver_type (components/science/source/solver/
!> sci_iterative_solver_mod.f90:38-57): the parent of the shim's device-resident solver types.
!> Only the abstract type is restated, so that the shim builds without the LFRic tree; the
!> concrete Krylov solvers are LFRic's own, and the drop-in test compiles them from the
!> reference sources.
module sci_iterative_solver_mod
  use constants_mod,           only : r_def, i_def, l_def
  use vector_mod,              only : abstract_vector_type
  use sci_linear_operator_mod, only : abstract_linear_operator_type
  use sci_preconditioner_mod,  only : abstract_preconditioner_type
  implicit none

  type, public, abstract :: abstract_iterative_solver_type
    private
    class(abstract_preconditioner_type),  pointer :: prec => null()
    class(abstract_linear_operator_type), pointer :: lin_op => null()
    real(kind=r_def)    :: r_tol
    real(kind=r_def)    :: a_tol
    integer(kind=i_def) :: max_iter
    logical(kind=l_def) :: monitor_convergence
    logical(kind=l_def) :: fail_on_non_converged
  contains
    procedure(apply_interface), deferred :: apply
  end type abstract_iterative_solver_type

  abstract interface
    subroutine apply_interface(self, x, b)
      import :: abstract_vector_type, abstract_iterative_solver_type
      class(abstract_iterative_solver_type), intent(inout) :: self
      class(abstract_vector_type),           intent(inout) :: x
      class(abstract_vector_type),           intent(inout) :: b
    end subroutine apply_interface
  end interface

end module sci_iterative_solver_mod
