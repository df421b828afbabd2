"""lfric_core_amd — MI355X-native LMA operator apply and Krylov solve for LFRic Core.

Python mirror of the reference's solver API over the C ABI of liblfric_hip.so
(include/lfric_hip.h).  Every call runs on the GPU through hand-written HIP kernels; there
is no CPU fallback: importing this package without the built library raises.

Reference interfaces mirrored (paths relative to the lfric_core tree):
  FieldVector            infrastructure/source/field/vector_mod.F90:17-59 (abstract_vector_type),
                         components/science/source/algorithm/solver/sci_field_vector_mod.x90,
                         sci_r_solver_field_vector_mod.x90 (FP32 storage, FP64 dots)
  LmaOperator            infrastructure/source/operator/operator_mod.t90 (operator_type) and
                         sci_mass_matrix_operator_alg_mod.x90:72-119 (apply = setval_c + matrix_vector)
  DiagonalPreconditioner sci_diagonal_preconditioner_alg_mod.x90:71-106
  NullPreconditioner     sci_null_preconditioner_alg_mod.x90
  ConjugateGradient      components/science/source/solver/sci_iterative_solver_mod.f90:87-121, 465-592
  GMRES, FGMRES, GCR,    the other abstract_iterative_solver_type extensions of
  BiCGStab, Jacobi,      sci_iterative_solver_mod.f90
  Chebyshev, PreconditionOnly
  solve                  the method factory of sci_mass_matrix_solver_alg_mod.x90:164-181
  ColumnwiseOperator     infrastructure/source/operator/columnwise_operator_mod.f90 and the
                         components/science/source/kernel/columnwise_operators/ kernels
  tri_solve              components/science/source/kernel/solver/sci_tri_solve_kernel_mod.f90
"""
from ._lib import (LfricError, NotConverged, lib, init, synchronize, timer_start, timer_stop,
                   device_bytes, set_apply_form, R64, R32, BF16, F16, W0, W2, W3)
from .fem import FunctionSpace, FieldVector, LmaOperator, extrusion_eta
from .cma import ColumnwiseOperator, ColumnwiseInversePreconditioner, tri_solve
from .solvers import (NullPreconditioner, DiagonalPreconditioner, ColumnJacobiPreconditioner,
                      LinearOperator, ConjugateGradient, MixedPrecisionCG, GMRES, MixedPrecisionGMRES,
                      FGMRES, GCR, BiCGStab, Jacobi, Chebyshev, PreconditionOnly, solve, METHODS,
                      SolverResult, SolverOptions)

__all__ = [
    "LfricError", "NotConverged", "lib", "init", "synchronize", "timer_start", "timer_stop",
    "device_bytes", "set_apply_form", "R64", "R32", "BF16", "F16", "W0", "W2", "W3", "FunctionSpace", "FieldVector",
    "LmaOperator", "NullPreconditioner", "DiagonalPreconditioner", "ColumnJacobiPreconditioner",
    "LinearOperator", "ConjugateGradient", "MixedPrecisionCG", "GMRES", "MixedPrecisionGMRES",
    "FGMRES", "GCR", "BiCGStab", "Jacobi", "Chebyshev", "PreconditionOnly", "solve", "METHODS",
    "SolverResult", "SolverOptions", "ColumnwiseOperator", "ColumnwiseInversePreconditioner", "tri_solve",
]
