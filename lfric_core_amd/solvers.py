"""Preconditioners and iterative solvers over the device-resident vectors and operators
(components/science/source/solver/sci_iterative_solver_mod.f90,
sci_preconditioner_mod.f90, components/science/source/algorithm/solver/)."""
import ctypes
from dataclasses import dataclass

from ._lib import (BF16, LFRIC_NOT_CONVERGED, LFRIC_OK, SolverOpts, SolverRes, check, lib)


class LinearOperator:
    """abstract_linear_operator_type (components/science/source/solver/
    sci_linear_operator_mod.f90:18-36): apply(x, y) computes y = A x on the device."""

    def __init__(self, handle, space, precision, keep=()):
        self._h, self.space, self.precision, self._keep = handle, space, precision, keep

    @classmethod
    def from_lma(cls, op):
        h = ctypes.c_void_p()
        check(lib().lfric_linop_create_lma(op._h, ctypes.byref(h)))
        return cls(h, op.fs_to, op.precision, (op,))

    @classmethod
    def mass_matrix(cls, op, apply_bc=True):
        """mass_matrix_operator_type(mass_matrix, apply_bc) (sci_mass_matrix_operator_alg_mod.x90:
        42-119): y = A x, then enforce_bc on y when apply_bc is set."""
        h = ctypes.c_void_p()
        check(lib().lfric_linop_create_mass_matrix(op._h, int(bool(apply_bc)), ctypes.byref(h)))
        return cls(h, op.fs_to, op.precision, (op,))

    @classmethod
    def helmholtz(cls, M3, D, m2_diag, tau):
        """y = M3 x + tau D diag(M2)^-1 D^T x on W3."""
        h = ctypes.c_void_p()
        check(lib().lfric_linop_create_helmholtz(M3._h, D._h, m2_diag._h, float(tau), ctypes.byref(h)))
        return cls(h, M3.fs_to, M3.precision, (M3, D, m2_diag))

    def apply(self, x, y):
        check(lib().lfric_linop_apply(self._h, x._h, y._h))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().lfric_linop_destroy(self._h)
            self._h = None


class ColumnJacobiPreconditioner:
    """Vertical-column block-Jacobi: tridiagonal column blocks of the assembled operator,
    factorised in FP32 and stored in `storage` precision (BF16 by default), applied in FP32."""

    def __init__(self, linop, storage=BF16):
        self.linop = linop
        self._h = ctypes.c_void_p()
        check(lib().lfric_prec_create_column_jacobi(linop._h, storage, ctypes.byref(self._h)))

    def apply(self, x, y):
        check(lib().lfric_prec_apply(self._h, x._h, y._h))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().lfric_prec_destroy(self._h)
            self._h = None


class NullPreconditioner:
    """null_preconditioner_type: y = x."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        check(lib().lfric_prec_create_identity(ctypes.byref(self._h)))

    def apply(self, x, y):
        check(lib().lfric_prec_apply(self._h, x._h, y._h))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().lfric_prec_destroy(self._h)
            self._h = None


class DiagonalPreconditioner:
    """diagonal_preconditioner_type (sci_diagonal_preconditioner_alg_mod.x90:71-106):
    y = x / diag, diag from mm_diagonal of the operator."""

    def __init__(self, diag):
        self.diag = diag  # keep the vector alive
        self._h = ctypes.c_void_p()
        check(lib().lfric_prec_create_diagonal(diag._h, ctypes.byref(self._h)))

    @classmethod
    def from_operator(cls, op):
        from .fem import FieldVector
        d = FieldVector(op.fs_to, op.precision)
        op.mm_diagonal(d)
        return cls(d)

    def apply(self, x, y):
        check(lib().lfric_prec_apply(self._h, x._h, y._h))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().lfric_prec_destroy(self._h)
            self._h = None


@dataclass
class SolverOptions:
    r_tol: float = 1e-8
    a_tol: float = 1e-16
    max_iter: int = 500
    gcrk: int = 30
    monitor_convergence: bool = True
    fail_on_non_converged: bool = False
    check_every: int = 8
    inner_r_tol: float = 1e-4
    max_outer: int = 20
    rho_relax: float = 1.0
    lmin: float = 0.124
    lmax: float = 3.377

    def c(self):
        return SolverOpts(self.r_tol, self.a_tol, self.max_iter, self.gcrk, int(self.monitor_convergence),
                          int(self.fail_on_non_converged), self.check_every, self.inner_r_tol, self.max_outer,
                          self.rho_relax, self.lmin, self.lmax)


@dataclass
class SolverResult:
    status: int
    iterations: int
    outer_iterations: int
    r_nrm_0: float
    r_nrm: float
    true_rel_res: float


class _Solver:
    def __init__(self, lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence=True,
                 fail_on_non_converged=False, **kw):
        # the reference stores pointers to lin_op and prec, not copies (sci_iterative_solver_mod.f90:478-479)
        self.lin_op, self.prec = lin_op, prec
        self.opts = SolverOptions(r_tol=r_tol, a_tol=a_tol, max_iter=max_iter,
                                  monitor_convergence=monitor_convergence,
                                  fail_on_non_converged=fail_on_non_converged, **kw)
        self.result = None

    def _finish(self, rc, res):
        self.result = SolverResult(res.status, res.iterations, res.outer_iterations, res.r_nrm_0, res.r_nrm,
                                   res.true_rel_res)
        check(rc, allow=(LFRIC_OK,) if self.opts.fail_on_non_converged else (LFRIC_OK, LFRIC_NOT_CONVERGED))
        return self.result


class ConjugateGradient(_Solver):
    """conjugate_gradient_type(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence,
    fail_on_non_converged) — apply(x, b) solves A x = b from the initial guess in x."""

    def apply(self, x, b):
        opts = self.opts.c()
        res = SolverRes()
        p = self.prec._h if self.prec is not None else None
        rc = lib().lfric_cg_solve(self.lin_op._h, p, x._h, b._h, ctypes.byref(opts), ctypes.byref(res))
        return self._finish(rc, res)


class MixedPrecisionCG(_Solver):
    """FP64 iterative refinement around FP32 (r_solver) CG.  lin_op is the FP64 operator,
    lin_op32 its FP32 copy, prec a preconditioner on FP32 vectors."""

    def __init__(self, lin_op, lin_op32, prec, r_tol, a_tol, max_iter, inner_r_tol=1e-4, **kw):
        super().__init__(lin_op, prec, r_tol, a_tol, max_iter, inner_r_tol=inner_r_tol, **kw)
        self.lin_op32 = lin_op32

    def apply(self, x, b):
        opts = self.opts.c()
        res = SolverRes()
        p = self.prec._h if self.prec is not None else None
        rc = lib().lfric_ir_cg_solve(self.lin_op._h, self.lin_op32._h, p, x._h, b._h, ctypes.byref(opts),
                                     ctypes.byref(res))
        return self._finish(rc, res)


def residual_norm(op, x, b):
    out = ctypes.c_double()
    check(lib().lfric_residual_norm(op._h, x._h, b._h, ctypes.byref(out)))
    return out.value


class GMRES(_Solver):
    """gmres_type(lin_op, prec, gcrk, r_tol, a_tol, max_iter, monitor_convergence,
    fail_on_non_converged) (sci_iterative_solver_mod.f90:161-196, 799-1003); max_iter counts
    restart cycles.  lin_op is a LinearOperator."""

    def __init__(self, lin_op, prec, gcrk, r_tol, a_tol, max_iter, monitor_convergence=True,
                 fail_on_non_converged=False, **kw):
        super().__init__(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence, fail_on_non_converged,
                         gcrk=gcrk, **kw)

    def apply(self, x, b):
        opts = self.opts.c()
        res = SolverRes()
        p = self.prec._h if self.prec is not None else None
        rc = lib().lfric_gmres_solve(self.lin_op._h, p, x._h, b._h, ctypes.byref(opts), ctypes.byref(res))
        return self._finish(rc, res)


class MixedPrecisionGMRES(_Solver):
    """FP64 iterative refinement around FP32 GMRES(gcrk); prec acts on FP32 vectors."""

    def __init__(self, lin_op, lin_op32, prec, gcrk, r_tol, a_tol, max_iter, inner_r_tol=1e-4, **kw):
        super().__init__(lin_op, prec, r_tol, a_tol, max_iter, gcrk=gcrk, inner_r_tol=inner_r_tol, **kw)
        self.lin_op32 = lin_op32

    def apply(self, x, b):
        opts = self.opts.c()
        res = SolverRes()
        p = self.prec._h if self.prec is not None else None
        rc = lib().lfric_ir_gmres_solve(self.lin_op._h, self.lin_op32._h, p, x._h, b._h, ctypes.byref(opts),
                                        ctypes.byref(res))
        return self._finish(rc, res)


METHOD_CG, METHOD_BICGSTAB, METHOD_GMRES, METHOD_FGMRES, METHOD_GCR, METHOD_PREC_ONLY, METHOD_JACOBI, \
    METHOD_CHEBYSHEV = range(1, 9)
METHODS = {"cg": METHOD_CG, "bicgstab": METHOD_BICGSTAB, "gmres": METHOD_GMRES, "fgmres": METHOD_FGMRES,
           "gcr": METHOD_GCR, "prec_only": METHOD_PREC_ONLY, "jacobi": METHOD_JACOBI, "chebyshev": METHOD_CHEBYSHEV}


class _LinopSolver(_Solver):
    _entry = None

    def apply(self, x, b):
        opts = self.opts.c()
        res = SolverRes()
        p = self.prec._h if self.prec is not None else None
        rc = getattr(lib(), self._entry)(self.lin_op._h, p, x._h, b._h, ctypes.byref(opts), ctypes.byref(res))
        return self._finish(rc, res)


class BiCGStab(_LinopSolver):
    """bicgstab_type(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence,
    fail_on_non_converged) (sci_iterative_solver_mod.f90:123-156, 645-779)."""
    _entry = "lfric_bicgstab_solve"


class FGMRES(_LinopSolver):
    """fgmres_type(lin_op, prec, gcrk, r_tol, a_tol, max_iter, ...) (sci_iterative_solver_mod.f90:
    199-233, 1055-1237); the residual starts from b (zero initial guess), max_iter counts
    restart cycles."""
    _entry = "lfric_fgmres_solve"

    def __init__(self, lin_op, prec, gcrk, r_tol, a_tol, max_iter, monitor_convergence=True,
                 fail_on_non_converged=False, **kw):
        super().__init__(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence, fail_on_non_converged,
                         gcrk=gcrk, **kw)


class GCR(FGMRES):
    """gcr_type(lin_op, prec, gcrk, r_tol, a_tol, max_iter, ...) (sci_iterative_solver_mod.f90:
    237-271, 1288-1408)."""
    _entry = "lfric_gcr_solve"


class Jacobi(_LinopSolver):
    """jacobi_type(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence,
    fail_on_non_converged, rho_relax) (sci_iterative_solver_mod.f90:350-387, jacobi_smod)."""
    _entry = "lfric_jacobi_solve"

    def __init__(self, lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence=True,
                 fail_on_non_converged=False, rho_relax=1.0, **kw):
        super().__init__(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence, fail_on_non_converged,
                         rho_relax=rho_relax, **kw)


class Chebyshev(_LinopSolver):
    """chebyshev_type(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence,
    fail_on_non_converged, lmin, lmax) (sci_iterative_solver_mod.f90:391-428, chebyshev_smod)."""
    _entry = "lfric_chebyshev_solve"

    def __init__(self, lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence=True,
                 fail_on_non_converged=False, lmin=0.124, lmax=3.377, **kw):
        super().__init__(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence, fail_on_non_converged,
                         lmin=lmin, lmax=lmax, **kw)


class PreconditionOnly(_LinopSolver):
    """precondition_only_type(lin_op, prec, monitor_convergence) (sci_iterative_solver_mod.f90:
    315-346, precondition_only_smod): x = P^-1 b."""
    _entry = "lfric_prec_only_solve"

    def __init__(self, lin_op, prec, monitor_convergence=True):
        super().__init__(lin_op, prec, 0.0, 0.0, 1, monitor_convergence)


def solve(method, lin_op, prec, x, b, **opts):
    """The solver factory of mass_matrix_solver_alg (sci_mass_matrix_solver_alg_mod.x90:164-181):
    method is one of METHODS' names; opts are SolverOptions fields."""
    o = SolverOptions(**opts).c()
    res = SolverRes()
    p = prec._h if prec is not None else None
    rc = lib().lfric_solve(METHODS[method], lin_op._h, p, x._h, b._h, ctypes.byref(o), ctypes.byref(res))
    out = SolverResult(res.status, res.iterations, res.outer_iterations, res.r_nrm_0, res.r_nrm,
                       res.true_rel_res)
    check(rc, allow=(LFRIC_OK,) if opts.get("fail_on_non_converged") else (LFRIC_OK, LFRIC_NOT_CONVERGED))
    return out
