"""Batched and served solves over the C ABI (lfric_batch_* in include/lfric_hip.h): many
independent mass-matrix solves, each with several right-hand sides, advanced together on the
device (FP32 CG inside FP64 iterative refinement per RHS)."""
import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import LFRIC_OK, R32, R64, W0, SolverRes, check, lib
from .solvers import SolverOptions, SolverResult


def batched_mix(n_problems, seed=0):
    """The batched configuration's problems: mesh C12/C24/C48, 10/30/70 levels, W0 element order
    0 or 1 (p = 0.5), 4 right-hand sides, a generator seed each, drawn from `seed`:
    [(edge cells, levels, order, nrhs, problem seed)]."""
    rng = np.random.default_rng(seed)
    return [(int(rng.choice((12, 24, 48))), int(rng.choice((10, 30, 70))), int(rng.integers(0, 2)), 4,
             int(rng.integers(0, 2**31))) for _ in range(n_problems)]


def prepare_batch_copies(op32):
    """The FP32 operator's copy for the batch's fastest apply: the row-gathered copy of a
    lowest-order operator, the symmetric-packed copy of an order-1 one."""
    if op32.fs_to.ndf == 8:
        op32.prepare_rows()
    elif op32.fs_to.ndf == 27:
        op32.prepare_symmetric()
    return op32


def build_problem(n, nl, order, nrhs, seed, scaled=False, copies=True, fp64_copies=False):
    """One batch problem on the device: the cubed-sphere W0 space, the FP64 operator
    (A_e = B_e B_e^T + 0.1 I from the counter-based generator) and its FP32 copy (with the
    batch copies of prepare_batch_copies when `copies`; on the FP64 operator too, for the
    refinement steps' residuals, when `fp64_copies`), the FP32 Jacobi diagonal, b ~ U(-1, 1)
    with nrhs columns and x = 0."""
    from .fem import FieldVector, FunctionSpace, LmaOperator
    fs = FunctionSpace.cubedsphere(n, nl, W0, order)
    op64 = LmaOperator(fs).generate_spd(seed, scaled=scaled)
    op32 = op64.convert(R32)
    if copies:
        prepare_batch_copies(op32)
    if fp64_copies:
        prepare_batch_copies(op64)
    d32 = op32.mm_diagonal(FieldVector(fs, R32))
    b = FieldVector(fs, R64, nrhs).fill_uniform(seed, 3)
    x = FieldVector(fs, R64, nrhs)
    x.set_scalar(0.0)
    return dict(fs=fs, op64=op64, op32=op32, d32=d32, b=b, x=x, nrhs=nrhs, key=(n, nl, order, nrhs, seed))


@dataclass
class BatchProblem:
    """One problem of a batch: FP64 and FP32 copies of a square LMA operator, its Jacobi
    diagonal (FP32) and x, b with nrhs right-hand sides (kept alive while in the batch)."""
    op64: object
    op32: object
    diag32: object
    x: object
    b: object


class Batch:
    def __init__(self, r_tol=1e-8, a_tol=0.0, max_iter=2000, check_every=8, inner_r_tol=1e-4, max_outer=20,
                 **kw):
        self.opts = SolverOptions(r_tol=r_tol, a_tol=a_tol, max_iter=max_iter, check_every=check_every,
                                  inner_r_tol=inner_r_tol, max_outer=max_outer, **kw)
        self._h = ctypes.c_void_p()
        o = self.opts.c()
        check(lib().lfric_batch_create(ctypes.byref(o), ctypes.byref(self._h)))
        self.problems = {}

    def add(self, op64, op32, diag32, x, b):
        assert diag32.precision == R32
        pid = ctypes.c_int()
        check(lib().lfric_batch_add(self._h, op64._h, op32._h, diag32._h, x._h, b._h, ctypes.byref(pid)))
        self.problems[pid.value] = BatchProblem(op64, op32, diag32, x, b)
        return pid.value

    def run(self, rounds=1):
        n = ctypes.c_int()
        check(lib().lfric_batch_run(self._h, int(rounds), ctypes.byref(n)))
        return n.value

    def solve(self):
        check(lib().lfric_batch_solve(self._h))

    def finished(self, pid):
        f = ctypes.c_int()
        check(lib().lfric_batch_finished(self._h, pid, ctypes.byref(f)))
        return bool(f.value)

    def result(self, pid):
        nr = self.problems[pid].x.nrhs
        res = (SolverRes * nr)()
        check(lib().lfric_batch_result(self._h, pid, res))
        return [SolverResult(r.status, r.iterations, r.outer_iterations, r.r_nrm_0, r.r_nrm, r.true_rel_res)
                for r in res]

    def remove(self, pid):
        check(lib().lfric_batch_remove(self._h, pid))
        del self.problems[pid]

    def tune(self):
        ms = ctypes.c_double()
        check(lib().lfric_batch_tune(self._h, ctypes.byref(ms)))
        return ms.value

    APPLY = {"gather": 0, "colour": 1, "copies": 2}

    def config(self, use_graph=-1, block=0, grid_mult=0, colour=None, apply=None):
        """use_graph 0/1, block 64/128/256, grid_mult > 0 (-1/0 keep).  apply: "copies"
        (default: an operator's row-gathered or symmetric-packed copy where it has one, else the
        colour-ordered forms), "colour" (colour-ordered forms where supported: lowest order on
        the matrix cores, order 1 register-staged) or "gather" (every problem in gather form);
        colour=True/False is "copies"/"gather"."""
        check(lib().lfric_batch_config(self._h, int(use_graph), int(block), int(grid_mult)))
        if colour is not None:
            apply = "copies" if colour else "gather"
        if apply is not None:
            check(lib().lfric_batch_set_apply(self._h, self.APPLY[apply]))

    def stats(self):
        it, blk, gm, nt = ctypes.c_longlong(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().lfric_batch_stats(self._h, ctypes.byref(it), ctypes.byref(blk), ctypes.byref(gm),
                                      ctypes.byref(nt)))
        return {"iterations": it.value, "block": blk.value, "grid_mult": gm.value, "tiles": nt.value}

    def close(self):
        if self._h:
            lib().lfric_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
