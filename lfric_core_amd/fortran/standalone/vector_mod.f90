!> abstract_vector_type of infrastructure/source/field/vector_mod.F90:17-59 (interfaces only).
module vector_mod
  use constants_mod, only : i_def, r_def
  implicit none
  private

  type, public, abstract :: abstract_vector_type
    private
  contains
    procedure(set_scalar_interface),  deferred :: set_scalar
    procedure(axpy_interface),        deferred :: axpy
    procedure(aypx_interface),        deferred :: aypx
    procedure(axpby_interface),       deferred :: axpby
    procedure(norm_interface),        deferred :: norm
    procedure(field_norm_interface),  deferred :: field_norm
    procedure(dot_interface),         deferred :: dot
    procedure(scale_interface),       deferred :: scale
    procedure(duplicate_interface),   deferred :: duplicate
    procedure(copy_interface),        deferred :: copy
    procedure(vector_size_interface), deferred :: vector_size
  end type abstract_vector_type

  abstract interface
    subroutine set_scalar_interface(self, scalar)
      import :: abstract_vector_type, r_def
      class(abstract_vector_type), intent(inout) :: self
      real(kind=r_def),            intent(in)    :: scalar
    end subroutine set_scalar_interface
    subroutine axpy_interface(self, alpha, x)
      import :: abstract_vector_type, r_def
      class(abstract_vector_type), intent(inout) :: self
      real(kind=r_def),            intent(in)    :: alpha
      class(abstract_vector_type), intent(inout) :: x
    end subroutine axpy_interface
    subroutine aypx_interface(self, alpha, x)
      import :: abstract_vector_type, r_def
      class(abstract_vector_type), intent(inout) :: self
      real(kind=r_def),            intent(in)    :: alpha
      class(abstract_vector_type), intent(inout) :: x
    end subroutine aypx_interface
    subroutine axpby_interface(self, a, x, b, y)
      import :: abstract_vector_type, r_def
      class(abstract_vector_type), intent(inout) :: self
      real(kind=r_def),            intent(in)    :: a, b
      class(abstract_vector_type), intent(in)    :: x, y
    end subroutine axpby_interface
    function norm_interface(self) result(normal)
      import :: abstract_vector_type, r_def
      class(abstract_vector_type), intent(in) :: self
      real(kind=r_def) :: normal
    end function norm_interface
    function field_norm_interface(self, n) result(normal)
      import :: abstract_vector_type, r_def, i_def
      class(abstract_vector_type), intent(in) :: self
      integer(kind=i_def),         intent(in) :: n
      real(kind=r_def) :: normal
    end function field_norm_interface
    function dot_interface(self, x) result(dot_prod)
      import :: abstract_vector_type, r_def
      class(abstract_vector_type), intent(in) :: self
      class(abstract_vector_type), intent(in) :: x
      real(kind=r_def) :: dot_prod
    end function dot_interface
    subroutine scale_interface(self, scalar)
      import :: abstract_vector_type, r_def
      class(abstract_vector_type), intent(inout) :: self
      real(kind=r_def),            intent(in)    :: scalar
    end subroutine scale_interface
    subroutine duplicate_interface(self, vec)
      import :: abstract_vector_type
      class(abstract_vector_type), intent(in)                 :: self
      class(abstract_vector_type), allocatable, intent(inout) :: vec
    end subroutine duplicate_interface
    subroutine copy_interface(self, source)
      import :: abstract_vector_type
      class(abstract_vector_type), intent(out) :: self
      class(abstract_vector_type), intent(in)  :: source
    end subroutine copy_interface
    function vector_size_interface(self) result(n_fields)
      import :: abstract_vector_type, i_def
      class(abstract_vector_type), intent(in) :: self
      integer(i_def) :: n_fields
    end function vector_size_interface
  end interface
end module vector_mod
