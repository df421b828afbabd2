!> Drop-in check of the HIP solver-API types: the stock conjugate_gradient_type and gmres_type
!> (written against the abstract types) run unmodified on hip_field_vector_type /
!> hip_lma_operator_type / hip_diagonal_preconditioner_type, and the device-resident
!> hip_conjugate_gradient_type gives the same solve; the stock bicgstab_type and the library's
!> BiCGStab behind hip_factory_solver_type (lfric_solve) agree.  Prints one JSON line; stops
!> with a non-zero code when the paths disagree.
program drop_in_test
  use iso_c_binding
  use constants_mod,            only : r_def, i_def
  use lfric_hip_capi_mod
  use lfric_hip_solver_api_mod
  use sci_iterative_solver_mod, only : conjugate_gradient_type, gmres_type, bicgstab_type
  implicit none
  type(hip_function_space_type) :: fs
  type(c_ptr) :: op
  type(hip_field_vector_type) :: b, diag, x_cg, x_hip, x_gm, x_bs, x_fbs
  type(hip_lma_operator_type) :: A
  type(hip_diagonal_preconditioner_type) :: P
  type(conjugate_gradient_type) :: cg
  type(gmres_type) :: gm
  type(hip_conjugate_gradient_type) :: hcg
  type(bicgstab_type) :: bs
  type(hip_factory_solver_type) :: fbs
  real(r_def), allocatable :: a1(:), a2(:), a3(:), a4(:), a5(:)
  real(r_def) :: d_cg, d_gm, d_bs, d_cma
  integer :: n, nl
  ! columnwise (CMA) path on W3: an operator assembled from an LMA W3 operator, its apply undone
  ! by its tridiagonal solve, and the solve used as a preconditioner
  type(hip_function_space_type) :: fs3
  type(c_ptr) :: op3, cma, pcma
  type(hip_field_vector_type) :: x3, y3, z3
  real(r_def), allocatable :: c1(:), c2(:)

  n = 12
  nl = 20
  call hip_check(lfric_init(0_c_int), "lfric_init")
  call fs%init_cubedsphere(n, nl, int(LFRIC_W0, i_def), 0)
  call hip_check(lfric_op_create(fs%h, fs%h, LFRIC_R64, op), "lfric_op_create")
  call hip_check(lfric_op_generate_spd(op, 0_c_long_long, 0_c_int, 0.1_c_double), "generate")
  A = hip_lma_operator(op)
  call diag%initialise(fs)
  call hip_check(lfric_mm_diagonal(op, diag%h), "mm_diagonal")
  P = hip_diagonal_preconditioner(diag)
  call b%initialise(fs)
  call hip_check(lfric_vec_fill_uniform(b%h, 0_c_long_long, 3_c_long_long), "rhs")

  call x_cg%initialise(fs)
  call x_cg%set_scalar(0.0_r_def)
  cg = conjugate_gradient_type(A, P, 1.0e-8_r_def, 1.0e-16_r_def, 500, .true., .true.)
  call cg%apply(x_cg, b)

  call x_hip%initialise(fs)
  call x_hip%set_scalar(0.0_r_def)
  hcg = hip_conjugate_gradient(A, P, 1.0e-8_r_def, 1.0e-16_r_def, 500, .true., .true.)
  call hcg%apply(x_hip, b)

  call x_gm%initialise(fs)
  call x_gm%set_scalar(0.0_r_def)
  gm = gmres_type(A, P, 10, 1.0e-10_r_def, 0.0_r_def, 50, .true., .true.)
  call gm%apply(x_gm, b)

  call x_bs%initialise(fs)
  call x_bs%set_scalar(0.0_r_def)
  bs = bicgstab_type(A, P, 1.0e-10_r_def, 1.0e-16_r_def, 300, .true., .true.)
  call bs%apply(x_bs, b)

  call x_fbs%initialise(fs)
  call x_fbs%set_scalar(0.0_r_def)
  fbs = hip_factory_solver(LFRIC_METHOD_BICGSTAB, A, P, 1.0e-10_r_def, 1.0e-16_r_def, 300, .true., .true.)
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
  write(*, '(A,I0,A,I0,A,I0,A,ES12.5,A,ES12.5,A,I0,A,I0,A,I0,A,ES12.5,A)') '{"undf": ', fs%undf, &
      ', "stock_cg_iterations": ', cg%iterations, ', "hip_cg_iterations": ', hcg%result%iterations, &
      ', "stock_vs_hip_cg_rel_diff": ', d_cg, ', "stock_gmres_vs_hip_cg_rel_diff": ', d_gm, &
      ', "gmres_cycles": ', gm%cycles, ', "stock_bicgstab_iterations": ', bs%iterations, &
      ', "hip_bicgstab_iterations": ', fbs%result%iterations, ', "stock_vs_hip_bicgstab_rel_diff": ', d_bs, '}'
  if (cg%iterations /= hcg%result%iterations .or. d_cg > 1.0e-10_r_def .or. d_gm > 1.0e-7_r_def) error stop 2
  if (bs%iterations /= fbs%result%iterations .or. d_bs > 1.0e-9_r_def) error stop 3

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
end program drop_in_test
