!> ISO_C_BINDING interfaces of the liblfric_hip C ABI (include/lfric_hip.h).
!>
!> The thin layer between LFRic's Fortran solver API and the HIP kernels: handles are
!> type(c_ptr), sizes are C ints, scalars C doubles, status codes C ints.
module lfric_hip_capi_mod
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: LFRIC_OK = 0, LFRIC_NOT_CONVERGED = 1, LFRIC_NAN = 2, &
                                       LFRIC_BREAKDOWN = 3
  integer(c_int), parameter, public :: LFRIC_R64 = 0, LFRIC_R32 = 1, LFRIC_BF16 = 2, LFRIC_F16 = 3
  integer(c_int), parameter, public :: LFRIC_W0 = 0, LFRIC_W2 = 2, LFRIC_W3 = 3
  ! solver methods of lfric_solve (the method enumeration of the mass-matrix solver factory)
  integer(c_int), parameter, public :: LFRIC_METHOD_CG = 1, LFRIC_METHOD_BICGSTAB = 2, &
                                       LFRIC_METHOD_GMRES = 3, LFRIC_METHOD_FGMRES = 4, &
                                       LFRIC_METHOD_GCR = 5, LFRIC_METHOD_PREC_ONLY = 6, &
                                       LFRIC_METHOD_JACOBI = 7, LFRIC_METHOD_CHEBYSHEV = 8

  type, bind(c), public :: lfric_solver_opts
    real(c_double) :: r_tol
    real(c_double) :: a_tol
    integer(c_int) :: max_iter
    integer(c_int) :: gcrk
    integer(c_int) :: monitor_convergence
    integer(c_int) :: fail_on_non_converged
    integer(c_int) :: check_every
    real(c_double) :: inner_r_tol
    integer(c_int) :: max_outer
    real(c_double) :: rho_relax
    real(c_double) :: lmin
    real(c_double) :: lmax
  end type lfric_solver_opts

  type, bind(c), public :: lfric_solver_result
    integer(c_int) :: status
    integer(c_int) :: iterations
    integer(c_int) :: outer_iterations
    real(c_double) :: r_nrm_0
    real(c_double) :: r_nrm
    real(c_double) :: true_rel_res
  end type lfric_solver_result

  public :: lfric_init, lfric_last_error_message, lfric_synchronize, lfric_device_bytes
  public :: lfric_fs_create, lfric_fs_create_cubedsphere, lfric_fs_info, lfric_fs_get_dofmap, &
            lfric_fs_get_colouring, lfric_fs_destroy
  public :: lfric_vec_create, lfric_vec_upload, lfric_vec_download, lfric_vec_destroy, &
            lfric_vec_upload_r32, lfric_vec_download_r32, &
            lfric_vec_set_scalar, lfric_vec_axpy, lfric_vec_aypx, lfric_vec_axpby, &
            lfric_vec_scale, lfric_vec_copy, lfric_vec_dot, lfric_vec_norm, lfric_vec_size, &
            lfric_vec_fill_uniform
  public :: lfric_op_create, lfric_op_upload, lfric_op_generate_spd, lfric_op_convert, &
            lfric_op_download, lfric_op_upload_r32, lfric_op_download_r32, &
            lfric_op_destroy, lfric_lma_apply, lfric_mm_diagonal, lfric_op_invert_local, &
            lfric_op_prepare_rows, lfric_op_tune, lfric_op_assemble_mass_w0, lfric_op_prepare_symmetric, &
            lfric_op_assemble_mass, lfric_op_assemble_div, lfric_enforce_bc, lfric_dg_inc_apply, &
            lfric_fs_get_boundary_dofs, lfric_fs_set_boundary_dofs
  public :: lfric_fs_create_ugrid
  public :: lfric_cma_create, lfric_cma_create_custom, lfric_cma_assemble, lfric_cma_assemble_diag_hmht, &
            lfric_cma_assemble_m2v_lumped_inv, lfric_cma_apply, lfric_cma_apply_inverse, lfric_cma_multiply, &
            lfric_cma_scaled_add, lfric_cma_destroy, lfric_prec_create_cma_inverse, lfric_tri_solve
  public :: lfric_prec_create_diagonal, lfric_prec_create_identity, lfric_prec_apply, &
            lfric_prec_destroy
  public :: lfric_linop_create_lma, lfric_linop_destroy, lfric_prec_create_column_jacobi
  public :: lfric_solver_opts_default, lfric_cg_solve, lfric_ir_cg_solve, lfric_gmres_solve, &
            lfric_solve

  interface
    integer(c_int) function lfric_init(device) bind(c, name="lfric_init")
      import :: c_int
      integer(c_int), value :: device
    end function
    type(c_ptr) function lfric_last_error_c() bind(c, name="lfric_last_error")
      import :: c_ptr
    end function
    integer(c_int) function lfric_synchronize() bind(c, name="lfric_synchronize")
      import :: c_int
    end function
    integer(c_long_long) function lfric_device_bytes() bind(c, name="lfric_device_bytes")
      import :: c_long_long
    end function

    integer(c_int) function lfric_fs_create(ncells, nlayers, ndf, undf, dofmap, ncolours, npc, cic, &
                                            maxcc, fs) bind(c, name="lfric_fs_create")
      import :: c_int, c_ptr
      integer(c_int), value :: ncells, nlayers, ndf, undf, ncolours, maxcc
      integer(c_int), intent(in) :: dofmap(*), npc(*), cic(*)
      type(c_ptr), intent(out) :: fs
    end function
    integer(c_int) function lfric_fs_create_cubedsphere(n, nlayers, space, order, fs) &
        bind(c, name="lfric_fs_create_cubedsphere")
      import :: c_int, c_ptr
      integer(c_int), value :: n, nlayers, space, order
      type(c_ptr), intent(out) :: fs
    end function
    integer(c_int) function lfric_fs_info(fs, ncells, nlayers, ndf, undf, ncolours) &
        bind(c, name="lfric_fs_info")
      import :: c_int, c_ptr
      type(c_ptr), value :: fs
      integer(c_int), intent(out) :: ncells, nlayers, ndf, undf, ncolours
    end function
    integer(c_int) function lfric_fs_get_dofmap(fs, dofmap) bind(c, name="lfric_fs_get_dofmap")
      import :: c_int, c_ptr
      type(c_ptr), value :: fs
      integer(c_int), intent(out) :: dofmap(*)
    end function
    integer(c_int) function lfric_fs_get_colouring(fs, npc, cic) bind(c, name="lfric_fs_get_colouring")
      import :: c_int, c_ptr
      type(c_ptr), value :: fs
      integer(c_int), intent(out) :: npc(*), cic(*)
    end function
    subroutine lfric_fs_destroy(fs) bind(c, name="lfric_fs_destroy")
      import :: c_ptr
      type(c_ptr), value :: fs
    end subroutine

    integer(c_int) function lfric_vec_create(fs, precision, nrhs, v) bind(c, name="lfric_vec_create")
      import :: c_int, c_ptr
      type(c_ptr), value :: fs
      integer(c_int), value :: precision, nrhs
      type(c_ptr), intent(out) :: v
    end function
    integer(c_int) function lfric_vec_upload(v, host) bind(c, name="lfric_vec_upload")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: v
      real(c_double), intent(in) :: host(*)
    end function
    integer(c_int) function lfric_vec_download(v, host) bind(c, name="lfric_vec_download")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: v
      real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function lfric_vec_upload_r32(v, host) bind(c, name="lfric_vec_upload_r32")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: v
      real(c_float), intent(in) :: host(*)
    end function
    integer(c_int) function lfric_vec_download_r32(v, host) bind(c, name="lfric_vec_download_r32")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: v
      real(c_float), intent(out) :: host(*)
    end function
    subroutine lfric_vec_destroy(v) bind(c, name="lfric_vec_destroy")
      import :: c_ptr
      type(c_ptr), value :: v
    end subroutine
    integer(c_int) function lfric_vec_size(v, undf, nrhs) bind(c, name="lfric_vec_size")
      import :: c_int, c_ptr
      type(c_ptr), value :: v
      integer(c_int), intent(out) :: undf, nrhs
    end function
    integer(c_int) function lfric_vec_set_scalar(v, s) bind(c, name="lfric_vec_set_scalar")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: v
      real(c_double), value :: s
    end function
    integer(c_int) function lfric_vec_axpy(y, a, x) bind(c, name="lfric_vec_axpy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: y, x
      real(c_double), value :: a
    end function
    integer(c_int) function lfric_vec_aypx(y, a, x) bind(c, name="lfric_vec_aypx")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: y, x
      real(c_double), value :: a
    end function
    integer(c_int) function lfric_vec_axpby(z, a, x, b, y) bind(c, name="lfric_vec_axpby")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: z, x, y
      real(c_double), value :: a, b
    end function
    integer(c_int) function lfric_vec_scale(v, a) bind(c, name="lfric_vec_scale")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: v
      real(c_double), value :: a
    end function
    integer(c_int) function lfric_vec_copy(dst, src) bind(c, name="lfric_vec_copy")
      import :: c_int, c_ptr
      type(c_ptr), value :: dst, src
    end function
    integer(c_int) function lfric_vec_dot(x, y, out) bind(c, name="lfric_vec_dot")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: x, y
      real(c_double), intent(out) :: out
    end function
    integer(c_int) function lfric_vec_norm(x, out) bind(c, name="lfric_vec_norm")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: x
      real(c_double), intent(out) :: out
    end function
    integer(c_int) function lfric_vec_fill_uniform(v, seed, stream) bind(c, name="lfric_vec_fill_uniform")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: v
      integer(c_long_long), value :: seed, stream
    end function

    integer(c_int) function lfric_op_create(to, from, precision, op) bind(c, name="lfric_op_create")
      import :: c_int, c_ptr
      type(c_ptr), value :: to, from
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: op
    end function
    integer(c_int) function lfric_op_upload(op, local_stencil) bind(c, name="lfric_op_upload")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: op
      real(c_double), intent(in) :: local_stencil(*)
    end function
    integer(c_int) function lfric_op_download(op, local_stencil) bind(c, name="lfric_op_download")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: op
      real(c_double), intent(out) :: local_stencil(*)
    end function
    integer(c_int) function lfric_op_upload_r32(op, local_stencil) bind(c, name="lfric_op_upload_r32")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: op
      real(c_float), intent(in) :: local_stencil(*)
    end function
    integer(c_int) function lfric_op_download_r32(op, local_stencil) bind(c, name="lfric_op_download_r32")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: op
      real(c_float), intent(out) :: local_stencil(*)
    end function
    integer(c_int) function lfric_op_generate_spd(op, seed, scaled, shift) bind(c, name="lfric_op_generate_spd")
      import :: c_int, c_ptr, c_long_long, c_double
      type(c_ptr), value :: op
      integer(c_long_long), value :: seed
      integer(c_int), value :: scaled
      real(c_double), value :: shift
    end function
    integer(c_int) function lfric_op_convert(src, dst) bind(c, name="lfric_op_convert")
      import :: c_int, c_ptr
      type(c_ptr), value :: src, dst
    end function
    ! invert_local_operator_code (sci_invert_local_operator_kernel_mod.F90:51-77)
    integer(c_int) function lfric_op_invert_local(a, ainv) bind(c, name="lfric_op_invert_local")
      import :: c_int, c_ptr
      type(c_ptr), value :: a, ainv
    end function
    ! compute_mass_matrix_w_scalar_code for lowest-order W0 from Cartesian chi fields
    integer(c_int) function lfric_op_assemble_mass_w0(op, chi1, chi2, chi3, nqp) &
        bind(c, name="lfric_op_assemble_mass_w0")
      import :: c_int, c_ptr
      type(c_ptr), value :: op, chi1, chi2, chi3
      integer(c_int), value :: nqp
    end function
    ! compute_mass_matrix_w2_code / _w3_code (or the W0 one) for the operator's lowest-order
    ! space, from Cartesian chi fields on a W0 space of the same mesh
    integer(c_int) function lfric_op_assemble_mass(op, chi1, chi2, chi3, nqp, rehabilitate) &
        bind(c, name="lfric_op_assemble_mass")
      import :: c_int, c_ptr
      type(c_ptr), value :: op, chi1, chi2, chi3
      integer(c_int), value :: nqp, rehabilitate
    end function
    ! a function space on the extruded mesh of a UGRID file; path null-terminated
    ! (trim(path)//c_null_char), mesh_name c_null_ptr for the file's only mesh
    integer(c_int) function lfric_fs_create_ugrid(path, mesh_name, nlayers, space, order, fs) &
        bind(c, name="lfric_fs_create_ugrid")
      import :: c_int, c_ptr, c_char
      character(kind=c_char), dimension(*), intent(in) :: path
      type(c_ptr), value :: mesh_name
      integer(c_int), value :: nlayers, space, order
      type(c_ptr), intent(out) :: fs
    end function
    ! columnwise_operator_type between horizontally discontinuous spaces (W3, Wtheta-like)
    integer(c_int) function lfric_cma_create(fs_to, fs_from, precision, a) bind(c, name="lfric_cma_create")
      import :: c_int, c_ptr
      type(c_ptr), value :: fs_to, fs_from
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: a
    end function
    integer(c_int) function lfric_cma_create_custom(fs_to, fs_from, precision, alpha, beta, gamma_m, gamma_p, a) &
        bind(c, name="lfric_cma_create_custom")
      import :: c_int, c_ptr
      type(c_ptr), value :: fs_to, fs_from
      integer(c_int), value :: precision, alpha, beta, gamma_m, gamma_p
      type(c_ptr), intent(out) :: a
    end function
    ! columnwise_op_asm_kernel_code from an LMA operator
    integer(c_int) function lfric_cma_assemble(a, lma) bind(c, name="lfric_cma_assemble")
      import :: c_int, c_ptr
      type(c_ptr), value :: a, lma
    end function
    integer(c_int) function lfric_cma_assemble_diag_hmht(a, dh, m2h) bind(c, name="lfric_cma_assemble_diag_hmht")
      import :: c_int, c_ptr
      type(c_ptr), value :: a, dh, m2h
    end function
    integer(c_int) function lfric_cma_assemble_m2v_lumped_inv(a, m) bind(c, name="lfric_cma_assemble_m2v_lumped_inv")
      import :: c_int, c_ptr
      type(c_ptr), value :: a, m
    end function
    ! columnwise_op_app_kernel_code (accumulate = 0: y = A x)
    integer(c_int) function lfric_cma_apply(a, x, y, accumulate) bind(c, name="lfric_cma_apply")
      import :: c_int, c_ptr
      type(c_ptr), value :: a, x, y
      integer(c_int), value :: accumulate
    end function
    ! columnwise_op_appinv_kernel_code
    integer(c_int) function lfric_cma_apply_inverse(a, x, y) bind(c, name="lfric_cma_apply_inverse")
      import :: c_int, c_ptr
      type(c_ptr), value :: a, x, y
    end function
    ! columnwise_op_mul_kernel_code: c = c + a b
    integer(c_int) function lfric_cma_multiply(a, b, c) bind(c, name="lfric_cma_multiply")
      import :: c_int, c_ptr
      type(c_ptr), value :: a, b, c
    end function
    ! columnwise_op_scaledadd_kernel_code: c = alpha a + beta b
    integer(c_int) function lfric_cma_scaled_add(a, b, c, alpha, beta) bind(c, name="lfric_cma_scaled_add")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: a, b, c
      real(c_double), value :: alpha, beta
    end function
    subroutine lfric_cma_destroy(a) bind(c, name="lfric_cma_destroy")
      import :: c_ptr
      type(c_ptr), value :: a
    end subroutine
    integer(c_int) function lfric_prec_create_cma_inverse(a, p) bind(c, name="lfric_prec_create_cma_inverse")
      import :: c_int, c_ptr
      type(c_ptr), value :: a
      type(c_ptr), intent(out) :: p
    end function
    ! tri_solve_code on W3 fields (work: a W3 scratch field)
    integer(c_int) function lfric_tri_solve(x, t0, tp, tm, y, work) bind(c, name="lfric_tri_solve")
      import :: c_int, c_ptr
      type(c_ptr), value :: x, t0, tp, tm, y, work
    end function
    ! compute_div_operator_code: weak divergence W2 -> W3
    integer(c_int) function lfric_op_assemble_div(op, chi1, chi2, chi3, nqp, rehabilitate) &
        bind(c, name="lfric_op_assemble_div")
      import :: c_int, c_ptr
      type(c_ptr), value :: op, chi1, chi2, chi3
      integer(c_int), value :: nqp, rehabilitate
    end function
    ! row-gathered copy for the one-launch W0 apply (set-up, once per operator)
    integer(c_int) function lfric_op_prepare_rows(op) bind(c, name="lfric_op_prepare_rows")
      import :: c_int, c_ptr
      type(c_ptr), value :: op
    end function
    ! symmetric-packed copy of a symmetric lowest-order W0 operator (0: 4 levels per thread)
    integer(c_int) function lfric_op_prepare_symmetric(op, levels_per_thread) &
        bind(c, name="lfric_op_prepare_symmetric")
      import :: c_int, c_ptr
      type(c_ptr), value :: op
      integer(c_int), value :: levels_per_thread
    end function
    ! apply autotuner (XCD block-order window), form 0 colour / 1 gather
    integer(c_int) function lfric_op_tune(op, form, best_window, best_ms) bind(c, name="lfric_op_tune")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: op
      integer(c_int), value :: form
      integer(c_int) :: best_window
      real(c_double) :: best_ms
    end function
    subroutine lfric_op_destroy(op) bind(c, name="lfric_op_destroy")
      import :: c_ptr
      type(c_ptr), value :: op
    end subroutine
    integer(c_int) function lfric_lma_apply(op, x, y) bind(c, name="lfric_lma_apply")
      import :: c_int, c_ptr
      type(c_ptr), value :: op, x, y
    end function
    integer(c_int) function lfric_dg_inc_apply(op, x, y) bind(c, name="lfric_dg_inc_apply")
      import :: c_int, c_ptr
      type(c_ptr), value :: op, x, y
    end function
    integer(c_int) function lfric_enforce_bc(v) bind(c, name="lfric_enforce_bc")
      import :: c_int, c_ptr
      type(c_ptr), value :: v
    end function
    integer(c_int) function lfric_fs_get_boundary_dofs(fs, bv) bind(c, name="lfric_fs_get_boundary_dofs")
      import :: c_int, c_ptr
      type(c_ptr), value :: fs
      integer(c_int), intent(out) :: bv(*)
    end function
    integer(c_int) function lfric_fs_set_boundary_dofs(fs, bv) bind(c, name="lfric_fs_set_boundary_dofs")
      import :: c_int, c_ptr
      type(c_ptr), value :: fs
      integer(c_int), intent(in) :: bv(*)
    end function
    integer(c_int) function lfric_mm_diagonal(op, diag) bind(c, name="lfric_mm_diagonal")
      import :: c_int, c_ptr
      type(c_ptr), value :: op, diag
    end function

    integer(c_int) function lfric_prec_create_diagonal(diag, p) bind(c, name="lfric_prec_create_diagonal")
      import :: c_int, c_ptr
      type(c_ptr), value :: diag
      type(c_ptr), intent(out) :: p
    end function
    integer(c_int) function lfric_prec_create_identity(p) bind(c, name="lfric_prec_create_identity")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: p
    end function
    integer(c_int) function lfric_prec_apply(p, x, y) bind(c, name="lfric_prec_apply")
      import :: c_int, c_ptr
      type(c_ptr), value :: p, x, y
    end function
    subroutine lfric_prec_destroy(p) bind(c, name="lfric_prec_destroy")
      import :: c_ptr
      type(c_ptr), value :: p
    end subroutine
    integer(c_int) function lfric_linop_create_lma(op, l) bind(c, name="lfric_linop_create_lma")
      import :: c_int, c_ptr
      type(c_ptr), value :: op
      type(c_ptr), intent(out) :: l
    end function
    subroutine lfric_linop_destroy(l) bind(c, name="lfric_linop_destroy")
      import :: c_ptr
      type(c_ptr), value :: l
    end subroutine
    integer(c_int) function lfric_prec_create_column_jacobi(l, storage, p) &
        bind(c, name="lfric_prec_create_column_jacobi")
      import :: c_int, c_ptr
      type(c_ptr), value :: l
      integer(c_int), value :: storage
      type(c_ptr), intent(out) :: p
    end function

    subroutine lfric_solver_opts_default(o) bind(c, name="lfric_solver_opts_default")
      import :: lfric_solver_opts
      type(lfric_solver_opts), intent(out) :: o
    end subroutine
    integer(c_int) function lfric_cg_solve(a, p, x, b, o, r) bind(c, name="lfric_cg_solve")
      import :: c_int, c_ptr, lfric_solver_opts, lfric_solver_result
      type(c_ptr), value :: a, p, x, b
      type(lfric_solver_opts), intent(in) :: o
      type(lfric_solver_result), intent(out) :: r
    end function
    integer(c_int) function lfric_ir_cg_solve(a64, a32, p32, x, b, o, r) bind(c, name="lfric_ir_cg_solve")
      import :: c_int, c_ptr, lfric_solver_opts, lfric_solver_result
      type(c_ptr), value :: a64, a32, p32, x, b
      type(lfric_solver_opts), intent(in) :: o
      type(lfric_solver_result), intent(out) :: r
    end function
    integer(c_int) function lfric_gmres_solve(a, p, x, b, o, r) bind(c, name="lfric_gmres_solve")
      import :: c_int, c_ptr, lfric_solver_opts, lfric_solver_result
      type(c_ptr), value :: a, p, x, b
      type(lfric_solver_opts), intent(in) :: o
      type(lfric_solver_result), intent(out) :: r
    end function
    integer(c_int) function lfric_solve(method, a, p, x, b, o, r) bind(c, name="lfric_solve")
      import :: c_int, c_ptr, lfric_solver_opts, lfric_solver_result
      integer(c_int), value :: method
      type(c_ptr), value :: a, p, x, b
      type(lfric_solver_opts), intent(in) :: o
      type(lfric_solver_result), intent(out) :: r
    end function
  end interface

  interface
    integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  !> Last error message of the library as a Fortran string.
  function lfric_last_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = lfric_last_error_c()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function lfric_last_error_message

end module lfric_hip_capi_mod
