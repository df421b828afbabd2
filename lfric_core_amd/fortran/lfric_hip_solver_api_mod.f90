!> HIP-resident implementations of LFRic's solver-API types.
!>
!>   hip_field_vector_type        extends abstract_vector_type (infrastructure/source/field/
!>                                vector_mod.F90:17-59); data stays in GPU memory (real64,
!>                                field_vector_type of sci_field_vector_mod.x90)
!>   hip_r_solver_field_vector_type
!>                                the same over real32 storage (r_solver_field_vector_type,
!>                                components/science/source/algorithm/solver/
!>                                sci_r_solver_field_vector_mod.x90:26-252): real(r_solver)
!>                                import/export, the real(r_def) scalars of axpy, aypx, axpby,
!>                                scale and set_scalar cast to r_solver before use (:189,213-214,
!>                                246), inner products accumulated in real64
!>   hip_lma_operator_type        extends abstract_linear_operator_type: y = A x with the
!>                                colour-ordered LMA kernels (replaces mass_matrix_operator_type%apply,
!>                                components/science/source/algorithm/solver/sci_mass_matrix_operator_alg_mod.x90:72-119)
!>   hip_diagonal_preconditioner_type
!>                                extends abstract_preconditioner_type (y = x / diag,
!>                                sci_diagonal_preconditioner_alg_mod.x90:71-106)
!>   hip_conjugate_gradient_type, hip_mixed_cg_type
!>                                extend abstract_iterative_solver_type and run the whole solve on
!>                                the device (FP64 CG; FP32 CG inside FP64 refinement)
!>
!> With these types LFRic's own conjugate_gradient_type / BiCGstab_type / gmres_type of
!> components/science/source/solver/sci_iterative_solver_mod.f90 run unmodified: every vector
!> operation they call becomes a HIP kernel.  Errors follow the reference: a failed library call
!> is logged at LOG_LEVEL_ERROR (which aborts), non-convergence at INFO or ERROR according to
!> fail_on_non_converged.
!>
!> Ownership: the reference duplicates vectors with sourced allocation
!> (allocate(v, source=x), sci_iterative_solver_mod.f90:883-885), which copies the device
!> handle.  Each object records its own address when it creates a device vector; only that
!> owner frees it, and a copy made by sourced allocation takes its own device vector before it
!> is first written.
module lfric_hip_solver_api_mod
  use iso_c_binding
  use constants_mod,           only : r_def, r_solver, i_def, l_def
  use log_mod,                 only : log_event, log_scratch_space, LOG_LEVEL_ERROR, LOG_LEVEL_INFO
  use vector_mod,              only : abstract_vector_type
  use sci_linear_operator_mod, only : abstract_linear_operator_type
  use sci_preconditioner_mod,  only : abstract_preconditioner_type
  use sci_iterative_solver_mod, only : abstract_iterative_solver_type
  use lfric_hip_capi_mod
  implicit none
  private

  public :: hip_check

  !> Function space handle (mesh, dofmap and colouring on the device)
  type, public :: hip_function_space_type
    type(c_ptr) :: h = c_null_ptr
    integer(i_def) :: ncells = 0, nlayers = 0, ndf = 0, undf = 0
  contains
    procedure :: init_cubedsphere => fs_init_cubedsphere
    procedure :: init_from_dofmap => fs_init_from_dofmap
  end type hip_function_space_type

  type, public, extends(abstract_vector_type) :: hip_field_vector_type
    type(c_ptr) :: h = c_null_ptr
    type(c_ptr) :: fs = c_null_ptr
    integer(c_int) :: precision = LFRIC_R64
    integer(c_int) :: undf = 0
    integer(c_intptr_t) :: owner = 0
  contains
    procedure :: initialise => vec_initialise
    procedure :: import_field => vec_import
    procedure :: export_field => vec_export
    procedure :: set_scalar => vec_set_scalar
    procedure :: axpy => vec_axpy
    procedure :: aypx => vec_aypx
    procedure :: axpby => vec_axpby
    procedure :: norm => vec_norm
    procedure :: field_norm => vec_field_norm
    procedure :: dot => vec_dot
    procedure :: scale => vec_scale
    procedure :: duplicate => vec_duplicate
    procedure :: copy => vec_copy
    procedure :: vector_size => vec_size
    procedure, private :: own
    final :: vec_final
  end type hip_field_vector_type

  type, public, extends(abstract_vector_type) :: hip_r_solver_field_vector_type
    type(c_ptr) :: h = c_null_ptr
    type(c_ptr) :: fs = c_null_ptr
    integer(c_int) :: undf = 0
    integer(c_intptr_t) :: owner = 0
  contains
    procedure :: initialise => rvec_initialise
    procedure :: import_field => rvec_import
    procedure :: export_field => rvec_export
    procedure :: set_scalar => rvec_set_scalar
    procedure :: axpy => rvec_axpy
    procedure :: aypx => rvec_aypx
    procedure :: axpby => rvec_axpby
    procedure :: norm => rvec_norm
    procedure :: field_norm => rvec_field_norm
    procedure :: dot => rvec_dot
    procedure :: scale => rvec_scale
    procedure :: duplicate => rvec_duplicate
    procedure :: copy => rvec_copy
    procedure :: vector_size => rvec_size
    procedure, private :: own => rvec_own
    final :: rvec_final
  end type hip_r_solver_field_vector_type

  !> y = A x over either vector type (the operator's storage precision must match the
  !> vectors'); `applies` counts the calls, so a caller can read how many products a solver made.
  !> With apply_bc set, y then takes the space's boundary conditions, as
  !> mass_matrix_operator_type%apply does (sci_mass_matrix_operator_alg_mod.x90:98-101).
  type, public, extends(abstract_linear_operator_type) :: hip_lma_operator_type
    type(c_ptr) :: op = c_null_ptr
    integer(i_def) :: applies = 0
    logical :: apply_bc = .false.
  contains
    procedure :: apply => op_apply
  end type hip_lma_operator_type

  type, public, extends(abstract_preconditioner_type) :: hip_diagonal_preconditioner_type
    type(c_ptr) :: p = c_null_ptr
  contains
    procedure :: apply => prec_apply
  end type hip_diagonal_preconditioner_type

  ! (the parent's private components lin_op, prec, r_tol, ... are not reachable from here, so
  ! the handles and options are held under their own names)
  type, public, extends(abstract_iterative_solver_type) :: hip_conjugate_gradient_type
    type(c_ptr) :: op_h = c_null_ptr, prec_h = c_null_ptr
    type(lfric_solver_opts) :: opts
    type(lfric_solver_result) :: result
  contains
    procedure :: apply => hip_cg_apply
  end type hip_conjugate_gradient_type

  type, public, extends(abstract_iterative_solver_type) :: hip_mixed_cg_type
    type(c_ptr) :: op64 = c_null_ptr, op32 = c_null_ptr, prec32 = c_null_ptr
    type(lfric_solver_opts) :: opts
    type(lfric_solver_result) :: result
  contains
    procedure :: apply => hip_mixed_cg_apply
  end type hip_mixed_cg_type

  ! any method of the mass-matrix solver factory, run by the library (lfric_solve)
  type, public, extends(abstract_iterative_solver_type) :: hip_factory_solver_type
    integer(c_int) :: method = LFRIC_METHOD_CG
    type(c_ptr) :: op_h = c_null_ptr, prec_h = c_null_ptr
    type(lfric_solver_opts) :: opts
    type(lfric_solver_result) :: result
  contains
    procedure :: apply => hip_factory_apply
  end type hip_factory_solver_type

  public :: hip_lma_operator, hip_diagonal_preconditioner, hip_conjugate_gradient, hip_mixed_cg, &
            hip_factory_solver

contains

  subroutine hip_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= LFRIC_OK) then
      write(log_scratch_space, '(A,": ",A," (status ",I0,")")') what, lfric_last_error_message(), rc
      call log_event(log_scratch_space, LOG_LEVEL_ERROR)
    end if
  end subroutine hip_check

  ! ---- function spaces ------------------------------------------------------------------------
  subroutine fs_fill(self)
    class(hip_function_space_type), intent(inout) :: self
    integer(c_int) :: nc, nl, ndf, undf, ncol
    call hip_check(lfric_fs_info(self%h, nc, nl, ndf, undf, ncol), "lfric_fs_info")
    self%ncells = nc
    self%nlayers = nl
    self%ndf = ndf
    self%undf = undf
  end subroutine fs_fill

  subroutine fs_init_cubedsphere(self, edge_cells, nlayers, space, order)
    class(hip_function_space_type), intent(inout) :: self
    integer(i_def), intent(in) :: edge_cells, nlayers, space, order
    call hip_check(lfric_fs_create_cubedsphere(edge_cells, nlayers, space, order, self%h), &
                   "lfric_fs_create_cubedsphere")
    call fs_fill(self)
  end subroutine fs_init_cubedsphere

  !> dofmap(ndf, ncells) and cells_in_colour(maxcc, ncolours) as LFRic's function_space_type
  !> and mesh_type hold them (1-based).
  subroutine fs_init_from_dofmap(self, nlayers, undf, dofmap, ncells_per_colour, cells_in_colour)
    class(hip_function_space_type), intent(inout) :: self
    integer(i_def), intent(in) :: nlayers, undf
    integer(i_def), intent(in) :: dofmap(:, :), ncells_per_colour(:), cells_in_colour(:, :)
    call hip_check(lfric_fs_create(size(dofmap, 2), nlayers, size(dofmap, 1), undf, dofmap, &
                                   size(ncells_per_colour), ncells_per_colour, cells_in_colour, &
                                   size(cells_in_colour, 1), self%h), "lfric_fs_create")
    call fs_fill(self)
  end subroutine fs_init_from_dofmap

  ! ---- vectors ----------------------------------------------------------------------------------
  subroutine vec_initialise(self, fs, precision)
    class(hip_field_vector_type), intent(inout) :: self
    type(hip_function_space_type), intent(in) :: fs
    integer(c_int), intent(in), optional :: precision
    self%fs = fs%h
    self%precision = LFRIC_R64
    if (present(precision)) self%precision = precision
    call hip_check(lfric_vec_create(self%fs, self%precision, 1_c_int, self%h), "lfric_vec_create")
    self%undf = fs%undf
    self%owner = loc(self)
  end subroutine vec_initialise

  !> Make sure this object owns its device vector (detach a copy made by sourced allocation).
  subroutine own(self)
    class(hip_field_vector_type), intent(inout) :: self
    type(c_ptr) :: shared
    if (self%owner == loc(self)) return
    shared = self%h
    call hip_check(lfric_vec_create(self%fs, self%precision, 1_c_int, self%h), "lfric_vec_create")
    if (c_associated(shared)) call hip_check(lfric_vec_copy(self%h, shared), "lfric_vec_copy")
    self%owner = loc(self)
  end subroutine own

  subroutine vec_final(self)
    type(hip_field_vector_type), intent(inout) :: self
    if (c_associated(self%h) .and. self%owner == loc(self)) call lfric_vec_destroy(self%h)
    self%h = c_null_ptr
  end subroutine vec_final

  subroutine vec_import(self, data)
    class(hip_field_vector_type), intent(inout) :: self
    real(r_def), intent(in) :: data(:)
    call self%own()
    call hip_check(lfric_vec_upload(self%h, data), "lfric_vec_upload")
  end subroutine vec_import

  subroutine vec_export(self, data)
    class(hip_field_vector_type), intent(in) :: self
    real(r_def), intent(out) :: data(:)
    call hip_check(lfric_vec_download(self%h, data), "lfric_vec_download")
  end subroutine vec_export

  subroutine vec_set_scalar(self, scalar)
    class(hip_field_vector_type), intent(inout) :: self
    real(kind=r_def), intent(in) :: scalar
    call self%own()
    call hip_check(lfric_vec_set_scalar(self%h, scalar), "set_scalar")
  end subroutine vec_set_scalar

  function handle_of(x, what) result(h)
    class(abstract_vector_type), intent(in) :: x
    character(len=*), intent(in) :: what
    type(c_ptr) :: h
    h = c_null_ptr
    select type (x)
    type is (hip_field_vector_type)
      h = x%h
    class default
      call log_event("hip_field_vector_type:" // what // ": argument is not a hip_field_vector_type", &
                     LOG_LEVEL_ERROR)
    end select
  end function handle_of

  subroutine vec_axpy(self, alpha, x)
    class(hip_field_vector_type), intent(inout) :: self
    real(kind=r_def), intent(in) :: alpha
    class(abstract_vector_type), intent(inout) :: x
    call self%own()
    call hip_check(lfric_vec_axpy(self%h, alpha, handle_of(x, "axpy")), "axpy")
  end subroutine vec_axpy

  subroutine vec_aypx(self, alpha, x)
    class(hip_field_vector_type), intent(inout) :: self
    real(kind=r_def), intent(in) :: alpha
    class(abstract_vector_type), intent(inout) :: x
    call self%own()
    call hip_check(lfric_vec_aypx(self%h, alpha, handle_of(x, "aypx")), "aypx")
  end subroutine vec_aypx

  subroutine vec_axpby(self, a, x, b, y)
    class(hip_field_vector_type), intent(inout) :: self
    real(kind=r_def), intent(in) :: a, b
    class(abstract_vector_type), intent(in) :: x, y
    call self%own()
    call hip_check(lfric_vec_axpby(self%h, a, handle_of(x, "axpby"), b, handle_of(y, "axpby")), "axpby")
  end subroutine vec_axpby

  function vec_norm(self) result(normal)
    class(hip_field_vector_type), intent(in) :: self
    real(kind=r_def) :: normal
    call hip_check(lfric_vec_norm(self%h, normal), "norm")
  end function vec_norm

  function vec_field_norm(self, n) result(normal)
    class(hip_field_vector_type), intent(in) :: self
    integer(kind=i_def), intent(in) :: n
    real(kind=r_def) :: normal
    if (n /= 1) call log_event("hip_field_vector_type holds a single field", LOG_LEVEL_ERROR)
    call hip_check(lfric_vec_norm(self%h, normal), "field_norm")
  end function vec_field_norm

  function vec_dot(self, x) result(dot_prod)
    class(hip_field_vector_type), intent(in) :: self
    class(abstract_vector_type), intent(in) :: x
    real(kind=r_def) :: dot_prod
    call hip_check(lfric_vec_dot(self%h, handle_of(x, "dot"), dot_prod), "dot")
  end function vec_dot

  subroutine vec_scale(self, scalar)
    class(hip_field_vector_type), intent(inout) :: self
    real(kind=r_def), intent(in) :: scalar
    call self%own()
    call hip_check(lfric_vec_scale(self%h, scalar), "scale")
  end subroutine vec_scale

  subroutine vec_duplicate(self, vec)
    class(hip_field_vector_type), intent(in) :: self
    class(abstract_vector_type), allocatable, intent(inout) :: vec
    if (allocated(vec)) deallocate(vec)
    allocate(hip_field_vector_type :: vec)
    select type (vec)
    type is (hip_field_vector_type)
      vec%fs = self%fs
      vec%precision = self%precision
      vec%undf = self%undf
      call hip_check(lfric_vec_create(vec%fs, vec%precision, 1_c_int, vec%h), "duplicate")
      vec%owner = loc(vec)
    end select
  end subroutine vec_duplicate

  !> self is intent(out), as copy_interface declares it (vector_mod.F90:173-177), so a vector
  !> passed here is finalised on entry (vec_final releases its device buffer to the library's
  !> cache) and this call takes a buffer of the source's shape.
  !> self is intent(out) because copy_interface declares it so (vector_mod.F90:173-177), and an
  !> override must keep the dummy's characteristics.  The processor finalizes the actual argument
  !> on entry (vec_final frees the handle it owns: F2008 4.5.6.3), and lfric_vec_create takes
  !> the freed buffer back from the library's vector cache, so the copies that the solvers make
  !> every iteration neither leak nor call hipMalloc/hipFree (the drop-in test checks the
  !> library's device bytes before and after a BiCGStab solve).
  subroutine vec_copy(self, source)
    class(hip_field_vector_type), intent(out) :: self
    class(abstract_vector_type), intent(in) :: source
    select type (source)
    type is (hip_field_vector_type)
      self%fs = source%fs
      self%precision = source%precision
      self%undf = source%undf
      call hip_check(lfric_vec_create(self%fs, self%precision, 1_c_int, self%h), "copy")
      self%owner = loc(self)
      call hip_check(lfric_vec_copy(self%h, source%h), "copy")
    class default
      call log_event("hip_field_vector_type:copy: source is not a hip_field_vector_type", LOG_LEVEL_ERROR)
    end select
  end subroutine vec_copy

  function vec_size(self) result(n_fields)
    class(hip_field_vector_type), intent(in) :: self
    integer(i_def) :: n_fields
    n_fields = 1
  end function vec_size

  ! ---- r_solver (real32) vectors ------------------------------------------------------------
  !> real(r_def) scalar as the reference passes it on: real(scalar, r_solver)
  !> (sci_r_solver_field_vector_mod.x90:189,213-214,246), then widened exactly for the C call
  elemental function to_rsol(a) result(b)
    real(r_def), intent(in) :: a
    real(r_def) :: b
    b = real(real(a, r_solver), r_def)
  end function to_rsol

  function rhandle_of(x, what) result(h)
    class(abstract_vector_type), intent(in) :: x
    character(len=*), intent(in) :: what
    type(c_ptr) :: h
    h = c_null_ptr
    select type (x)
    type is (hip_r_solver_field_vector_type)
      h = x%h
    class default
      call log_event("hip_r_solver_field_vector_type:" // what // &
                     ": argument is not a hip_r_solver_field_vector_type", LOG_LEVEL_ERROR)
    end select
  end function rhandle_of

  subroutine rvec_initialise(self, fs)
    class(hip_r_solver_field_vector_type), intent(inout) :: self
    type(hip_function_space_type), intent(in) :: fs
    self%fs = fs%h
    call hip_check(lfric_vec_create(self%fs, LFRIC_R32, 1_c_int, self%h), "lfric_vec_create")
    self%undf = fs%undf
    self%owner = loc(self)
  end subroutine rvec_initialise

  subroutine rvec_own(self)
    class(hip_r_solver_field_vector_type), intent(inout) :: self
    type(c_ptr) :: shared
    if (self%owner == loc(self)) return
    shared = self%h
    call hip_check(lfric_vec_create(self%fs, LFRIC_R32, 1_c_int, self%h), "lfric_vec_create")
    if (c_associated(shared)) call hip_check(lfric_vec_copy(self%h, shared), "lfric_vec_copy")
    self%owner = loc(self)
  end subroutine rvec_own

  subroutine rvec_final(self)
    type(hip_r_solver_field_vector_type), intent(inout) :: self
    if (c_associated(self%h) .and. self%owner == loc(self)) call lfric_vec_destroy(self%h)
    self%h = c_null_ptr
  end subroutine rvec_final

  subroutine rvec_import(self, data)
    class(hip_r_solver_field_vector_type), intent(inout) :: self
    real(r_solver), intent(in) :: data(:)
    call self%own()
    call hip_check(lfric_vec_upload_r32(self%h, data), "lfric_vec_upload_r32")
  end subroutine rvec_import

  subroutine rvec_export(self, data)
    class(hip_r_solver_field_vector_type), intent(in) :: self
    real(r_solver), intent(out) :: data(:)
    call hip_check(lfric_vec_download_r32(self%h, data), "lfric_vec_download_r32")
  end subroutine rvec_export

  subroutine rvec_set_scalar(self, scalar)
    class(hip_r_solver_field_vector_type), intent(inout) :: self
    real(kind=r_def), intent(in) :: scalar
    call self%own()
    call hip_check(lfric_vec_set_scalar(self%h, to_rsol(scalar)), "set_scalar")
  end subroutine rvec_set_scalar

  subroutine rvec_axpy(self, alpha, x)
    class(hip_r_solver_field_vector_type), intent(inout) :: self
    real(kind=r_def), intent(in) :: alpha
    class(abstract_vector_type), intent(inout) :: x
    call self%own()
    call hip_check(lfric_vec_axpy(self%h, to_rsol(alpha), rhandle_of(x, "axpy")), "axpy")
  end subroutine rvec_axpy

  subroutine rvec_aypx(self, alpha, x)
    class(hip_r_solver_field_vector_type), intent(inout) :: self
    real(kind=r_def), intent(in) :: alpha
    class(abstract_vector_type), intent(inout) :: x
    call self%own()
    call hip_check(lfric_vec_aypx(self%h, to_rsol(alpha), rhandle_of(x, "aypx")), "aypx")
  end subroutine rvec_aypx

  subroutine rvec_axpby(self, a, x, b, y)
    class(hip_r_solver_field_vector_type), intent(inout) :: self
    real(kind=r_def), intent(in) :: a, b
    class(abstract_vector_type), intent(in) :: x, y
    call self%own()
    call hip_check(lfric_vec_axpby(self%h, to_rsol(a), rhandle_of(x, "axpby"), to_rsol(b), &
                                   rhandle_of(y, "axpby")), "axpby")
  end subroutine rvec_axpby

  function rvec_norm(self) result(normal)
    class(hip_r_solver_field_vector_type), intent(in) :: self
    real(kind=r_def) :: normal
    call hip_check(lfric_vec_norm(self%h, normal), "norm")
  end function rvec_norm

  function rvec_field_norm(self, n) result(normal)
    class(hip_r_solver_field_vector_type), intent(in) :: self
    integer(kind=i_def), intent(in) :: n
    real(kind=r_def) :: normal
    if (n /= 1) call log_event("hip_r_solver_field_vector_type holds a single field", LOG_LEVEL_ERROR)
    call hip_check(lfric_vec_norm(self%h, normal), "field_norm")
  end function rvec_field_norm

  function rvec_dot(self, x) result(dot_prod)
    class(hip_r_solver_field_vector_type), intent(in) :: self
    class(abstract_vector_type), intent(in) :: x
    real(kind=r_def) :: dot_prod
    call hip_check(lfric_vec_dot(self%h, rhandle_of(x, "dot"), dot_prod), "dot")
  end function rvec_dot

  subroutine rvec_scale(self, scalar)
    class(hip_r_solver_field_vector_type), intent(inout) :: self
    real(kind=r_def), intent(in) :: scalar
    call self%own()
    call hip_check(lfric_vec_scale(self%h, to_rsol(scalar)), "scale")
  end subroutine rvec_scale

  subroutine rvec_duplicate(self, vec)
    class(hip_r_solver_field_vector_type), intent(in) :: self
    class(abstract_vector_type), allocatable, intent(inout) :: vec
    if (allocated(vec)) deallocate(vec)
    allocate(hip_r_solver_field_vector_type :: vec)
    select type (vec)
    type is (hip_r_solver_field_vector_type)
      vec%fs = self%fs
      vec%undf = self%undf
      call hip_check(lfric_vec_create(vec%fs, LFRIC_R32, 1_c_int, vec%h), "duplicate")
      vec%owner = loc(vec)
    end select
  end subroutine rvec_duplicate

  !> intent(out) as copy_interface declares it; see vec_copy
  !> intent(out) as copy_interface requires; see vec_copy for the handle's life
  subroutine rvec_copy(self, source)
    class(hip_r_solver_field_vector_type), intent(out) :: self
    class(abstract_vector_type), intent(in) :: source
    select type (source)
    type is (hip_r_solver_field_vector_type)
      self%fs = source%fs
      self%undf = source%undf
      call hip_check(lfric_vec_create(self%fs, LFRIC_R32, 1_c_int, self%h), "copy")
      self%owner = loc(self)
      call hip_check(lfric_vec_copy(self%h, source%h), "copy")
    class default
      call log_event("hip_r_solver_field_vector_type:copy: source is not a hip_r_solver_field_vector_type", &
                     LOG_LEVEL_ERROR)
    end select
  end subroutine rvec_copy

  function rvec_size(self) result(n_fields)
    class(hip_r_solver_field_vector_type), intent(in) :: self
    integer(i_def) :: n_fields
    n_fields = 1
  end function rvec_size

  !> device handle of a vector of either kind (operators and preconditioners take both)
  function any_handle(x, what) result(h)
    class(abstract_vector_type), intent(in) :: x
    character(len=*), intent(in) :: what
    type(c_ptr) :: h
    h = c_null_ptr
    select type (x)
    type is (hip_field_vector_type)
      h = x%h
    type is (hip_r_solver_field_vector_type)
      h = x%h
    class default
      call log_event(what // ": argument is not a HIP field vector", LOG_LEVEL_ERROR)
    end select
  end function any_handle

  ! ---- operator and preconditioner ------------------------------------------------------------
  function hip_lma_operator(op) result(self)
    type(c_ptr), intent(in) :: op
    type(hip_lma_operator_type) :: self
    self%op = op
  end function hip_lma_operator

  subroutine op_apply(self, x, y)
    class(hip_lma_operator_type), intent(inout) :: self
    class(abstract_vector_type), intent(in) :: x
    class(abstract_vector_type), intent(inout) :: y
    select type (y)
    type is (hip_field_vector_type)
      call y%own()
      call hip_check(lfric_lma_apply(self%op, any_handle(x, "lma_apply"), y%h), "lma_apply")
      if (self%apply_bc) call hip_check(lfric_enforce_bc(y%h), "enforce_bc")
    type is (hip_r_solver_field_vector_type)
      call y%own()
      call hip_check(lfric_lma_apply(self%op, any_handle(x, "lma_apply"), y%h), "lma_apply")
      if (self%apply_bc) call hip_check(lfric_enforce_bc(y%h), "enforce_bc")
    class default
      call log_event("hip_lma_operator_type:apply: y is not a HIP field vector", LOG_LEVEL_ERROR)
    end select
    self%applies = self%applies + 1
  end subroutine op_apply

  !> diag is a hip_field_vector_type or a hip_r_solver_field_vector_type; the preconditioner
  !> then divides vectors of that kind
  function hip_diagonal_preconditioner(diag) result(self)
    class(abstract_vector_type), intent(in) :: diag
    type(hip_diagonal_preconditioner_type) :: self
    call hip_check(lfric_prec_create_diagonal(any_handle(diag, "diagonal preconditioner"), self%p), &
                   "lfric_prec_create_diagonal")
  end function hip_diagonal_preconditioner

  subroutine prec_apply(self, x, y)
    class(hip_diagonal_preconditioner_type), intent(inout) :: self
    class(abstract_vector_type), intent(in) :: x
    class(abstract_vector_type), intent(inout) :: y
    select type (y)
    type is (hip_field_vector_type)
      call y%own()
      call hip_check(lfric_prec_apply(self%p, any_handle(x, "prec"), y%h), "prec_apply")
    type is (hip_r_solver_field_vector_type)
      call y%own()
      call hip_check(lfric_prec_apply(self%p, any_handle(x, "prec"), y%h), "prec_apply")
    class default
      call log_event("hip_diagonal_preconditioner_type: y is not a HIP field vector", LOG_LEVEL_ERROR)
    end select
  end subroutine prec_apply

  ! ---- device-resident solvers ----------------------------------------------------------------
  function hip_conjugate_gradient(op, prec, r_tol, a_tol, max_iter, monitor_convergence, &
                                  fail_on_non_converged) result(self)
    type(hip_lma_operator_type), intent(in) :: op
    type(hip_diagonal_preconditioner_type), intent(in) :: prec
    real(r_def), intent(in) :: r_tol, a_tol
    integer(i_def), intent(in) :: max_iter
    logical(l_def), intent(in) :: monitor_convergence, fail_on_non_converged
    type(hip_conjugate_gradient_type) :: self
    self%op_h = op%op
    self%prec_h = prec%p
    call lfric_solver_opts_default(self%opts)
    self%opts%r_tol = r_tol
    self%opts%a_tol = a_tol
    self%opts%max_iter = max_iter
    self%opts%monitor_convergence = merge(1, 0, monitor_convergence)
    self%opts%fail_on_non_converged = merge(1, 0, fail_on_non_converged)
  end function hip_conjugate_gradient

  subroutine report(rc, res, name)
    integer(c_int), intent(in) :: rc
    type(lfric_solver_result), intent(in) :: res
    character(len=*), intent(in) :: name
    if (rc == LFRIC_NOT_CONVERGED) then
      write(log_scratch_space, '(A," failed to converge after ",I6," iterations")') name, res%iterations
      call log_event(log_scratch_space, LOG_LEVEL_ERROR)
    else if (rc /= LFRIC_OK) then
      call hip_check(rc, name)
    else if (res%status == LFRIC_NOT_CONVERGED) then
      write(log_scratch_space, '(A," failed to converge after ",I6," iterations")') name, res%iterations
      call log_event(log_scratch_space, LOG_LEVEL_INFO)
    end if
  end subroutine report

  subroutine hip_cg_apply(self, x, b)
    class(hip_conjugate_gradient_type), intent(inout) :: self
    class(abstract_vector_type), intent(inout) :: x
    class(abstract_vector_type), intent(inout) :: b
    integer(c_int) :: rc
    select type (x)
    type is (hip_field_vector_type)
      call x%own()
      rc = lfric_cg_solve(self%op_h, self%prec_h, x%h, handle_of(b, "cg"), self%opts, self%result)
      call report(rc, self%result, "hip cg")
    class default
      call log_event("hip_conjugate_gradient_type: x is not a hip_field_vector_type", LOG_LEVEL_ERROR)
    end select
  end subroutine hip_cg_apply

  function hip_mixed_cg(op64, op32, prec32, r_tol, max_iter, inner_r_tol) result(self)
    type(c_ptr), intent(in) :: op64, op32, prec32
    real(r_def), intent(in) :: r_tol, inner_r_tol
    integer(i_def), intent(in) :: max_iter
    type(hip_mixed_cg_type) :: self
    self%op64 = op64
    self%op32 = op32
    self%prec32 = prec32
    call lfric_solver_opts_default(self%opts)
    self%opts%r_tol = r_tol
    self%opts%a_tol = 0.0_c_double
    self%opts%max_iter = max_iter
    self%opts%inner_r_tol = inner_r_tol
  end function hip_mixed_cg

  subroutine hip_mixed_cg_apply(self, x, b)
    class(hip_mixed_cg_type), intent(inout) :: self
    class(abstract_vector_type), intent(inout) :: x
    class(abstract_vector_type), intent(inout) :: b
    integer(c_int) :: rc
    select type (x)
    type is (hip_field_vector_type)
      call x%own()
      rc = lfric_ir_cg_solve(self%op64, self%op32, self%prec32, x%h, handle_of(b, "mixed cg"), self%opts, &
                             self%result)
      call report(rc, self%result, "hip mixed-precision cg")
    class default
      call log_event("hip_mixed_cg_type: x is not a hip_field_vector_type", LOG_LEVEL_ERROR)
    end select
  end subroutine hip_mixed_cg_apply

  !> The solver of sci_mass_matrix_solver_alg_mod.x90:164-283 for `method` (LFRIC_METHOD_*),
  !> with the options of its constructor there; gcrk for GMRES / FGMRES / GCR, rho_relax for
  !> Jacobi, lmin / lmax for Chebyshev.
  function hip_factory_solver(method, op, prec, r_tol, a_tol, max_iter, monitor_convergence, &
                              fail_on_non_converged, gcrk, rho_relax, lmin, lmax) result(self)
    integer(c_int), intent(in) :: method
    type(hip_lma_operator_type), intent(in) :: op
    type(hip_diagonal_preconditioner_type), intent(in) :: prec
    real(r_def), intent(in) :: r_tol, a_tol
    integer(i_def), intent(in) :: max_iter
    logical(l_def), intent(in) :: monitor_convergence, fail_on_non_converged
    integer(i_def), intent(in), optional :: gcrk
    real(r_def), intent(in), optional :: rho_relax, lmin, lmax
    type(hip_factory_solver_type) :: self
    self%method = method
    self%op_h = op%op
    self%prec_h = prec%p
    call lfric_solver_opts_default(self%opts)
    self%opts%r_tol = r_tol
    self%opts%a_tol = a_tol
    self%opts%max_iter = max_iter
    self%opts%monitor_convergence = merge(1, 0, monitor_convergence)
    self%opts%fail_on_non_converged = merge(1, 0, fail_on_non_converged)
    if (present(gcrk)) self%opts%gcrk = gcrk
    if (present(rho_relax)) self%opts%rho_relax = rho_relax
    if (present(lmin)) self%opts%lmin = lmin
    if (present(lmax)) self%opts%lmax = lmax
  end function hip_factory_solver

  subroutine hip_factory_apply(self, x, b)
    class(hip_factory_solver_type), intent(inout) :: self
    class(abstract_vector_type), intent(inout) :: x
    class(abstract_vector_type), intent(inout) :: b
    integer(c_int) :: rc
    type(c_ptr) :: linop
    select type (x)
    type is (hip_field_vector_type)
      call x%own()
      ! the abstract linear operator over the LMA operator lives for this solve only, so copies
      ! of the solver object own nothing
      call hip_check(lfric_linop_create_lma(self%op_h, linop), "linop")
      rc = lfric_solve(self%method, linop, self%prec_h, x%h, handle_of(b, "solve"), self%opts, self%result)
      call lfric_linop_destroy(linop)
      call report(rc, self%result, "hip solver")
    class default
      call log_event("hip_factory_solver_type: x is not a hip_field_vector_type", LOG_LEVEL_ERROR)
    end select
  end subroutine hip_factory_apply

end module lfric_hip_solver_api_mod
