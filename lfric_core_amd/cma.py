"""Columnwise (CMA) operators and the vertical tridiagonal solve on the GPU.

Reference interfaces mirrored (paths relative to the lfric_core tree):
  ColumnwiseOperator  infrastructure/source/operator/columnwise_operator_mod.f90
                      (columnwise_operator_type: alpha, beta, gamma_m, gamma_p, bandwidth, nrow,
                      ncol) and the kernels of components/science/source/kernel/
                      columnwise_operators/ (asm, asm_diag_hmht, asm_m2v_lumped_inv, app,
                      appinv, mul, scaledadd)
  tri_solve           components/science/source/kernel/solver/sci_tri_solve_kernel_mod.f90
"""
import ctypes

import numpy as np

from ._lib import R32, check, init, lib, ptr


class ColumnwiseOperator:
    """One banded matrix per cell column between two horizontally discontinuous spaces of the
    same mesh (W3, Wtheta-like), held on the device."""

    def __init__(self, fs_to, fs_from=None, precision=R32, custom=None):
        init()
        self.fs_to = fs_to
        self.fs_from = fs_from or fs_to
        self.precision = precision
        self._h = ctypes.c_void_p()
        if custom is None:
            check(lib().lfric_cma_create(self.fs_to._h, self.fs_from._h, precision, ctypes.byref(self._h)))
        else:
            check(lib().lfric_cma_create_custom(self.fs_to._h, self.fs_from._h, precision, *[int(v) for v in custom],
                                                ctypes.byref(self._h)))
        p = np.zeros(7, np.int32)
        check(lib().lfric_cma_info(self._h, ptr(p)))
        self.params = p  # alpha, beta, gamma_m, gamma_p, bandwidth, nrow, ncol

    @property
    def bandwidth(self):
        return int(self.params[4])

    @property
    def nrow(self):
        return int(self.params[5])

    @property
    def ncol(self):
        return int(self.params[6])

    def upload(self, ref):
        """ref: columnwise_matrix as a C-order (ncell_2d, nrow, bandwidth) array (the reference's
        Fortran (bandwidth, nrow, ncell_2d) byte for byte)."""
        a = np.ascontiguousarray(ref, np.float64)
        assert a.shape == (self.fs_to.ncells, self.nrow, self.bandwidth)
        check(lib().lfric_cma_upload(self._h, ptr(a)))
        return self

    def download(self):
        out = np.zeros((self.fs_to.ncells, self.nrow, self.bandwidth))
        check(lib().lfric_cma_download(self._h, ptr(out)))
        return out

    def assemble(self, lma):
        """columnwise_op_asm of an LMA operator on the same spaces."""
        check(lib().lfric_cma_assemble(self._h, lma._h))
        return self

    def assemble_diag_hmht(self, Dh, M2h):
        check(lib().lfric_cma_assemble_diag_hmht(self._h, Dh._h, M2h._h))
        return self

    def assemble_m2v_lumped_inv(self, M):
        check(lib().lfric_cma_assemble_m2v_lumped_inv(self._h, M._h))
        return self

    def apply(self, x, y, accumulate=False):
        """y = A x (accumulate: y += A x)."""
        check(lib().lfric_cma_apply(self._h, x._h, y._h, int(bool(accumulate))))

    def apply_inverse(self, x, y):
        """y = A^-1 x, tridiagonal solve of every column."""
        check(lib().lfric_cma_apply_inverse(self._h, x._h, y._h))

    def multiply_into(self, B, C):
        """C += self B."""
        check(lib().lfric_cma_multiply(self._h, B._h, C._h))
        return C

    def scaled_add_into(self, B, C, alpha, beta):
        """C = alpha self + beta B."""
        check(lib().lfric_cma_scaled_add(self._h, B._h, C._h, float(alpha), float(beta)))
        return C

    def __del__(self):
        if getattr(self, "_h", None):
            lib().lfric_cma_destroy(self._h)
            self._h = None


class ColumnwiseInversePreconditioner:
    """y = A^-1 x with A a square tridiagonal ColumnwiseOperator: the vertical line solve as a
    preconditioner of the Krylov solvers."""

    def __init__(self, cma):
        self.cma = cma  # keep the operator alive
        self._h = ctypes.c_void_p()
        check(lib().lfric_prec_create_cma_inverse(cma._h, ctypes.byref(self._h)))

    def apply(self, x, y):
        check(lib().lfric_prec_apply(self._h, x._h, y._h))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().lfric_prec_destroy(self._h)
            self._h = None


def tri_solve(x, t0, tp, tm, y, work):
    """tri_solve_code on every W3 column: y = T^-1 x, T with diagonal t0, super-diagonal tp and
    sub-diagonal tm; work is a scratch field of the same space and precision."""
    check(lib().lfric_tri_solve(x._h, t0._h, tp._h, tm._h, y._h, work._h))
