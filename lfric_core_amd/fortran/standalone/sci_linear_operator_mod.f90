!> abstract_linear_operator_type of components/science/source/solver/sci_linear_operator_mod.f90:18-36.
module sci_linear_operator_mod
  use vector_mod, only : abstract_vector_type
  implicit none
  private
  type, public, abstract :: abstract_linear_operator_type
    private
  contains
    procedure(apply_interface), deferred :: apply
  end type abstract_linear_operator_type
  abstract interface
    subroutine apply_interface(self, x, y)
      import :: abstract_linear_operator_type, abstract_vector_type
      class(abstract_linear_operator_type), intent(inout) :: self
      class(abstract_vector_type),          intent(in)    :: x
      class(abstract_vector_type),          intent(inout) :: y
    end subroutine apply_interface
  end interface
end module sci_linear_operator_mod
