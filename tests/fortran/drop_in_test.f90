!> Drop-in check of the HIP solver-API types against LFRic's own solver module.
!>
!> Built by oracle/Makefile.dropin against the reference's sci_iterative_solver_mod.f90,
!> vector_mod.F90, sci_preconditioner_mod.f90, log_mod.F90 and constants_mod.F90 compiled from
!> the LFRic sources, so conjugate_gradient_type, BiCGstab_type and gmres_type below are LFRic's
!> code driving hip_field_vector_type / hip_r_solver_field_vector_type, hip_lma_operator_type
!> and hip_diagonal_preconditioner_type.  The solvers log through LFRic's log_mod ("cg converged
!> after N iterations", ...); the operator wrappers count their applies.  Prints JSON lines and
!> stops with a non-zero code when a check fails; the pytest wrapper compares the iteration
!> counts with the oracle.
program drop_in_test
  use iso_c_binding
  use constants_mod,            only : r_def, r_solver, i_def
  use log_mod,                  only : initialise_logging
  use lfric_hip_capi_mod
  use lfric_hip_solver_api_mod
  use sci_iterative_solver_mod, only : conjugate_gradient_type, gmres_type, BiCGstab_type
  implicit none
  type(hip_function_space_type) :: fs
  type(c_ptr) :: op, op32
  type(hip_field_vector_type) :: b, diag, x_cg, x_hip, x_gm, x_bs, x_fbs
  type(hip_r_solver_field_vector_type) :: b32, diag32, x32
  type(hip_lma_operator_type), target :: A_cg, A_gm, A_bs, A_fbs, A32
  type(hip_diagonal_preconditioner_type), target :: P, P32
  type(hip_conjugate_gradient_type) :: hcg
  type(hip_factory_solver_type) :: fbs
  real(r_def), allocatable :: a1(:), a2(:), a3(:), a4(:), a5(:), stencil(:), bh(:)
  real(r_solver), allocatable :: stencil32(:), back32(:), b32h(:), x32h(:)
  real(r_def) :: d_cg, d_gm, d_bs, d_cma
  integer(c_long_long) :: bytes_before, bytes_after
  integer :: n, nl, cg_applies, bs_applies, ncell3d
  ! columnwise (CMA) path on W3: an operator assembled from an LMA W3 operator, its apply undone
  ! by its tridiagonal solve, and the solve used as a preconditioner
  type(hip_function_space_type) :: fs3
  type(c_ptr) :: op3, cma, pcma
  type(hip_field_vector_type) :: x3, y3, z3
  real(r_def), allocatable :: c1(:), c2(:)

  call initialise_logging(0, "drop_in_test")
  n = 12
  nl = 20
  call hip_check(lfric_init(0_c_int), "lfric_init")
  call fs%init_cubedsphere(n, nl, int(LFRIC_W0, i_def), 0)
  call hip_check(lfric_op_create(fs%h, fs%h, LFRIC_R64, op), "lfric_op_create")
  call hip_check(lfric_op_generate_spd(op, 0_c_long_long, 0_c_int, 0.1_c_double), "generate")
  A_cg = hip_lma_operator(op)
  A_gm = hip_lma_operator(op)
  A_bs = hip_lma_operator(op)
  A_fbs = hip_lma_operator(op)
  call diag%initialise(fs)
  call hip_check(lfric_mm_diagonal(op, diag%h), "mm_diagonal")
  P = hip_diagonal_preconditioner(diag)
  call b%initialise(fs)
  call hip_check(lfric_vec_fill_uniform(b%h, 0_c_long_long, 3_c_long_long), "rhs")

  ! LFRic's conjugate_gradient_type on real64 HIP vectors, and the device-resident solve
  call x_cg%initialise(fs)
  call x_cg%set_scalar(0.0_r_def)
  call run_cg(A_cg, P, x_cg, b, 1.0e-8_r_def, 1.0e-16_r_def)
  cg_applies = A_cg%applies

  call x_hip%initialise(fs)
  call x_hip%set_scalar(0.0_r_def)
  hcg = hip_conjugate_gradient(A_cg, P, 1.0e-8_r_def, 1.0e-16_r_def, 500, .true., .true.)
  call hcg%apply(x_hip, b)

  ! LFRic's gmres_type
  call x_gm%initialise(fs)
  call x_gm%set_scalar(0.0_r_def)
  call run_gmres(A_gm, P, x_gm, b)

  ! LFRic's BiCGstab_type (copy() several times per iteration: the device bytes held by live
  ! handles must be the same before and after), and the library's BiCGStab
  call x_bs%initialise(fs)
  call x_bs%set_scalar(0.0_r_def)
  bytes_before = lfric_device_bytes()
  call run_bicgstab(A_bs, P, x_bs, b)
  bytes_after = lfric_device_bytes()
  bs_applies = A_bs%applies

  call x_fbs%initialise(fs)
  call x_fbs%set_scalar(0.0_r_def)
  fbs = hip_factory_solver(LFRIC_METHOD_BICGSTAB, A_fbs, P, 1.0e-10_r_def, 1.0e-16_r_def, 300, .true., .true.)
  call fbs%apply(x_fbs, b)

  allocate(a1(fs%undf), a2(fs%undf), a3(fs%undf), a4(fs%undf), a5(fs%undf))
  call x_cg%export_field(a1)
  call x_hip%export_field(a2)
  call x_gm%export_field(a3)
  call x_bs%export_field(a4)
  call x_fbs%export_field(a5)
  d_cg = norm2(a1 - a2) / norm2(a2)
  d_gm = norm2(a3 - a2) / norm2(a2)
  d_bs = norm2(a4 - a5) / norm2(a5)
  write(*, '(A,I0,A,I0,A,I0,A,ES12.5,A,ES12.5,A,I0,A,I0,A,ES12.5,A,I0,A,I0,A)') '{"undf": ', fs%undf, &
      ', "stock_cg_applies": ', cg_applies, ', "hip_cg_iterations": ', hcg%result%iterations, &
      ', "stock_vs_hip_cg_rel_diff": ', d_cg, ', "stock_gmres_vs_hip_cg_rel_diff": ', d_gm, &
      ', "stock_bicgstab_applies": ', bs_applies, ', "hip_bicgstab_iterations": ', fbs%result%iterations, &
      ', "stock_vs_hip_bicgstab_rel_diff": ', d_bs, ', "device_bytes_before_bicgstab": ', bytes_before, &
      ', "device_bytes_after_bicgstab": ', bytes_after, '}'
  if (cg_applies - 1 /= hcg%result%iterations .or. d_cg > 1.0e-10_r_def .or. d_gm > 1.0e-7_r_def) error stop 2
  if (d_bs > 1.0e-9_r_def .or. bytes_before /= bytes_after) error stop 3

  ! r_solver: the same operator rounded to real32 on the host, uploaded as
  ! r_solver_operator_type data, and LFRic's conjugate_gradient_type on real32 HIP vectors
  ncell3d = fs%ncells * fs%nlayers
  allocate(stencil(ncell3d * fs%ndf * fs%ndf), stencil32(ncell3d * fs%ndf * fs%ndf), &
           back32(ncell3d * fs%ndf * fs%ndf))
  call hip_check(lfric_op_download(op, stencil), "lfric_op_download")
  stencil32 = real(stencil, r_solver)
  call hip_check(lfric_op_create(fs%h, fs%h, LFRIC_R32, op32), "lfric_op_create r32")
  call hip_check(lfric_op_upload_r32(op32, stencil32), "lfric_op_upload_r32")
  call hip_check(lfric_op_download_r32(op32, back32), "lfric_op_download_r32")
  if (any(back32 /= stencil32)) error stop 5
  A32 = hip_lma_operator(op32)
  call diag32%initialise(fs)
  call hip_check(lfric_mm_diagonal(op32, diag32%h), "mm_diagonal r32")
  P32 = hip_diagonal_preconditioner(diag32)
  allocate(bh(fs%undf), b32h(fs%undf), x32h(fs%undf))
  call b%export_field(bh)
  b32h = real(bh, r_solver)
  call b32%initialise(fs)
  call b32%import_field(b32h)
  call x32%initialise(fs)
  call x32%set_scalar(0.0_r_def)
  call run_cg(A32, P32, x32, b32, 1.0e-5_r_def, 0.0_r_def)
  call x32%export_field(x32h)
  write(*, '(A,I0,A,ES12.5,A,ES12.5,A)') '{"r_solver_cg_applies": ', A32%applies, &
      ', "r_solver_vs_hip_cg_rel_diff": ', norm2(real(x32h, r_def) - a2) / norm2(a2), &
      ', "r_solver_x_norm": ', norm2(real(x32h, r_def)), '}'

  ! columnwise operators from Fortran
  call fs3%init_cubedsphere(n, nl, int(LFRIC_W3, i_def), 0)
  call hip_check(lfric_op_create(fs3%h, fs3%h, LFRIC_R64, op3), "lfric_op_create W3")
  call hip_check(lfric_op_generate_spd(op3, 1_c_long_long, 0_c_int, 0.5_c_double), "generate W3")
  call hip_check(lfric_cma_create_custom(fs3%h, fs3%h, LFRIC_R64, 1_c_int, 1_c_int, 1_c_int, 1_c_int, cma), &
                 "lfric_cma_create_custom")
  call hip_check(lfric_cma_assemble(cma, op3), "lfric_cma_assemble")
  call x3%initialise(fs3)
  call y3%initialise(fs3)
  call z3%initialise(fs3)
  call hip_check(lfric_vec_fill_uniform(x3%h, 4_c_long_long, 3_c_long_long), "x3")
  call hip_check(lfric_cma_apply(cma, x3%h, y3%h, 0_c_int), "lfric_cma_apply")
  call hip_check(lfric_cma_apply_inverse(cma, y3%h, z3%h), "lfric_cma_apply_inverse")
  allocate(c1(fs3%undf), c2(fs3%undf))
  call x3%export_field(c1)
  call z3%export_field(c2)
  d_cma = norm2(c2 - c1) / norm2(c1)
  call hip_check(lfric_prec_create_cma_inverse(cma, pcma), "lfric_prec_create_cma_inverse")
  call hip_check(lfric_prec_apply(pcma, y3%h, z3%h), "lfric_prec_apply")
  call z3%export_field(c2)
  write(*, '(A,ES12.5,A,ES12.5,A)') '{"cma_solve_undoes_apply_rel_diff": ', d_cma, &
      ', "cma_preconditioner_rel_diff": ', norm2(c2 - c1) / norm2(c1), '}'
  if (d_cma > 1.0e-13_r_def .or. norm2(c2 - c1) / norm2(c1) > 1.0e-13_r_def) error stop 4
  call lfric_prec_destroy(pcma)
  call lfric_cma_destroy(cma)

contains

  ! Each solver lives in its own scope, so its temporaries are released when it returns, as in an
  ! LFRic algorithm.
  subroutine run_cg(A, P, x, b, r_tol, a_tol)
    type(hip_lma_operator_type), intent(inout), target :: A
    type(hip_diagonal_preconditioner_type), intent(inout), target :: P
    class(*), intent(inout) :: x, b
    real(r_def), intent(in) :: r_tol, a_tol
    type(conjugate_gradient_type) :: cg
    cg = conjugate_gradient_type(A, P, r_tol, a_tol, 500, .true., .true.)
    select type (x)
    type is (hip_field_vector_type)
      select type (b)
      type is (hip_field_vector_type)
        call cg%apply(x, b)
      end select
    type is (hip_r_solver_field_vector_type)
      select type (b)
      type is (hip_r_solver_field_vector_type)
        call cg%apply(x, b)
      end select
    end select
  end subroutine run_cg

  subroutine run_gmres(A, P, x, b)
    type(hip_lma_operator_type), intent(inout), target :: A
    type(hip_diagonal_preconditioner_type), intent(inout), target :: P
    type(hip_field_vector_type), intent(inout) :: x, b
    type(gmres_type) :: gm
    gm = gmres_type(A, P, 10, 1.0e-10_r_def, 0.0_r_def, 50, .true., .true.)
    call gm%apply(x, b)
  end subroutine run_gmres

  subroutine run_bicgstab(A, P, x, b)
    type(hip_lma_operator_type), intent(inout), target :: A
    type(hip_diagonal_preconditioner_type), intent(inout), target :: P
    type(hip_field_vector_type), intent(inout) :: x, b
    type(BiCGstab_type) :: bs
    bs = BiCGstab_type(A, P, 1.0e-10_r_def, 1.0e-16_r_def, 300, .true., .true.)
    call bs%apply(x, b)
  end subroutine run_bicgstab

end program drop_in_test
