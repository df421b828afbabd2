"""Batched and served solves over the C ABI (lfric_batch_* in include/lfric_hip.h): many
independent mass-matrix solves, each with several right-hand sides, advanced together on the
device (FP32 CG inside FP64 iterative refinement per RHS)."""
import ctypes
from dataclasses import dataclass

from ._lib import LFRIC_OK, R32, SolverRes, check, lib
from .solvers import SolverOptions, SolverResult


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

    def config(self, use_graph=-1, block=0, grid_mult=0, colour=None):
        """use_graph 0/1, block 64/128/256, grid_mult > 0 (-1/0 keep); colour True/False: the
        W0 problems apply in colour-ordered form (lowest order on the matrix cores; default) or
        every problem in gather form."""
        check(lib().lfric_batch_config(self._h, int(use_graph), int(block), int(grid_mult)))
        if colour is not None:
            check(lib().lfric_batch_set_apply(self._h, int(bool(colour))))

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
