"""The allocation-free CG entry point against lfric_cg_solve, through the C ABI.

lfric_cg_solve_noalloc runs the same kernels as lfric_cg_solve with its vectors and partials in
the library's static device scratch array, so its solution and iteration count must equal the
allocating entry point's bit for bit, on two calls in a row (the second on a smaller problem,
so that stale contents of the array would show), and no device memory may be allocated."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve(hip, entry, op, prec, b, precision):
    from lfric_core_amd import _lib
    from lfric_core_amd.solvers import SolverOptions, SolverRes
    x = hip.FieldVector(op.fs_to, precision)
    x.set_scalar(0.0)
    opts = SolverOptions(r_tol=1e-8, a_tol=1e-16, max_iter=500).c()
    res = SolverRes()
    before = _lib.lib().lfric_device_bytes()
    rc = getattr(_lib.lib(), entry)(op._h, prec._h, x._h, b._h, ctypes.byref(opts), ctypes.byref(res))
    after = _lib.lib().lfric_device_bytes()
    assert rc == 0, (entry, rc, _lib.lib().lfric_last_error())
    return x.export_raw(), res.iterations, res.r_nrm, after - before


def _problem(hip, n, nl, precision):
    space = hip.FunctionSpace.cubedsphere(n, nl)
    op = hip.LmaOperator(space).generate_spd(0)
    if precision == hip.R32:
        op = op.convert(hip.R32)
    diag = hip.FieldVector(space, precision)
    op.mm_diagonal(diag)
    prec = hip.DiagonalPreconditioner(diag)
    b = hip.FieldVector(space, precision)
    b.fill_uniform(7, 3)
    return space, op, prec, b


@pytest.mark.parametrize("precision", ["R64", "R32"])
def test_noalloc_cg_equals_cg_on_two_calls(hip, precision):
    prec_kind = getattr(hip, precision)
    for n, nl in ((12, 20), (6, 8)):  # the second call is smaller
        print(f"noalloc CG: C{n}x{nl} {precision}", flush=True)
        _, op, prec, b = _problem(hip, n, nl, prec_kind)
        x_ref, it_ref, nrm_ref, _ = _solve(hip, "lfric_cg_solve", op, prec, b, prec_kind)
        x_na, it_na, nrm_na, grown = _solve(hip, "lfric_cg_solve_noalloc", op, prec, b, prec_kind)
        assert grown == 0, "the allocation-free CG allocated device memory"
        assert it_na == it_ref and it_ref > 5
        assert nrm_na == nrm_ref
        assert np.array_equal(x_na.view(np.uint8), x_ref.view(np.uint8)), "solutions differ"


def test_noalloc_cg_refusals(hip):
    from lfric_core_amd import _lib
    from lfric_core_amd.solvers import SolverOptions, SolverRes
    L = _lib.lib()
    opts = SolverOptions().c()
    res = SolverRes()
    # a fresh space: its gather lists do not exist yet, and the entry point does not build them
    space = hip.FunctionSpace.cubedsphere(5, 4)
    op = hip.LmaOperator(space).generate_spd(1)
    x, b = hip.FieldVector(space), hip.FieldVector(space)
    b.fill_uniform(1, 3)
    x.set_scalar(0.0)
    rc = L.lfric_cg_solve_noalloc(op._h, None, x._h, b._h, ctypes.byref(opts), ctypes.byref(res))
    assert rc == -4 and b"gather lists" in L.lfric_last_error()
    y = hip.FieldVector(space)
    op.apply_gather(b, y)  # builds them
    rc = L.lfric_cg_solve_noalloc(op._h, None, x._h, b._h, ctypes.byref(opts), ctypes.byref(res))
    assert rc == 0 and res.iterations > 0
    # beyond the 64 MiB maximum: three FP64 vectors of C84x70 W0 (undf = 3,006,098)
    big = hip.FunctionSpace.cubedsphere(84, 70)
    opb = hip.LmaOperator(big)
    xb, bb = hip.FieldVector(big), hip.FieldVector(big)
    rc = L.lfric_cg_solve_noalloc(opb._h, None, xb._h, bb._h, ctypes.byref(opts), ctypes.byref(res))
    assert rc == -5 and b"maximum" in L.lfric_last_error()
