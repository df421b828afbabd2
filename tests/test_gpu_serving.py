"""GPU test of the served stream: continuous batching under an HBM budget.

Bar: every request finishes with true FP64 relative residual <= its tolerance, and its
solution is bit-identical to the same request solved alone (its own batch, nothing else on
the device)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone(hip):
    from lfric_core_amd import serving
    from lfric_core_amd.batch import Batch
    reqs = serving.poisson_stream(300.0, 0.06, seed=3, meshes=(4, 6), weights=(0.6, 0.4), levels=(3, 5))
    assert len(reqs) >= 8
    # a budget that holds about three requests at once forces waiting and backfill
    budget = 3 * max(serving.footprint_bytes(r.n, r.nl, r.order, r.nrhs) for r in reqs)
    srv = serving.SolveServer(budget, starve_s=0.05, check_every=4)
    keep = {}
    orig_live = serving._Live

    class Recording(orig_live):
        def __init__(self, req):
            super().__init__(req)
            keep[req.rid] = self

    serving._Live = Recording
    try:
        rep = srv.serve(reqs, wall_limit_s=30.0)
    finally:
        serving._Live = orig_live
        srv.close()
    assert rep.completed == len(reqs) and rep.rejected == 0
    assert rep.peak_bytes <= budget
    for r in reqs:
        assert r.status == 0 and r.true_rel_res <= r.tol, (r.rid, r.status, r.true_rel_res)
        lv = keep[r.rid]
        mixed = lv.x.export_field().copy()
        alone = serving._Live(r)
        bt = Batch(r_tol=r.tol, a_tol=0.0, check_every=4, inner_r_tol=1e-4, max_iter=4000)
        pid = bt.add(alone.op64, alone.op32, alone.diag32, alone.x, alone.b)
        bt.solve()
        np.testing.assert_array_equal(alone.x.export_field(), mixed)
        bt.remove(pid)
        bt.close()
