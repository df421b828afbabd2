!> abstract_preconditioner_type of components/science/source/solver/sci_preconditioner_mod.f90:19-38.
module sci_preconditioner_mod
  use vector_mod, only : abstract_vector_type
  implicit none
  private
  type, public, abstract :: abstract_preconditioner_type
    private
  contains
    procedure(apply_interface), deferred :: apply
  end type abstract_preconditioner_type
  abstract interface
    subroutine apply_interface(self, x, y)
      import :: abstract_vector_type, abstract_preconditioner_type
      class(abstract_preconditioner_type), intent(inout) :: self
      class(abstract_vector_type),         intent(in)    :: x
      class(abstract_vector_type),         intent(inout) :: y
    end subroutine apply_interface
  end interface
end module sci_preconditioner_mod
