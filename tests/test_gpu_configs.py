"""The batched and served configurations at their own workloads on the GPU.

Batched: the 64 uneven problems of the batched configuration (C12/C24/C48, 10/30/70 levels,
order 0/1, 4 right-hand sides each, seed 0) in one batch: every right-hand side reaches an FP64
true relative residual <= 1e-8 (computed on the device by the refinement step), and three
sampled problems are bit-identical when solved alone.

Served: the first 12 requests, in arrival order, of the served configuration's stream (Poisson
200 requests/s, seed 0: C24-C96 here, 30/70 levels, order 0/1, 1-8 RHS, tolerances 1e-6/1e-8,
scaled element matrices) through the continuous-batching server under the device's HBM budget:
every request meets its tolerance, and three sampled requests are bit-identical alone."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batched_configuration_64_problems(hip):
    from lfric_core_amd.batch import Batch, batched_mix, build_problem
    mix = batched_mix(64, seed=0)
    assert len(mix) == 64 and {m[0] for m in mix} == {12, 24, 48} and {m[2] for m in mix} == {0, 1}
    t0 = time.perf_counter()
    probs = [build_problem(*k, fp64_copies=True) for k in mix]
    print(f"built 64 problems in {time.perf_counter() - t0:.1f} s", flush=True)
    bt = Batch(r_tol=1e-8, a_tol=0.0, max_iter=4000, check_every=8)
    ids = [bt.add(p["op64"], p["op32"], p["d32"], p["x"], p["b"]) for p in probs]
    t0 = time.perf_counter()
    bt.solve()
    print(f"batch solved in {time.perf_counter() - t0:.2f} s", flush=True)
    mixed = {}
    for i, pid in enumerate(ids):
        assert bt.finished(pid)
        res = bt.result(pid)
        assert len(res) == 4
        for r in res:
            assert r.status == 0 and r.true_rel_res <= 1e-8, (i, mix[i], r)
        mixed[i] = probs[i]["x"].export_field().copy()
    bt.close()
    # three sampled problems (one of each mesh where the mix allows), solved alone
    sample = []
    for n in (12, 24, 48):
        sample.append(next(i for i, m in enumerate(mix) if m[0] == n))
    for i in sample:
        p = probs[i]
        p["x"].set_scalar(0.0)
        alone = Batch(r_tol=1e-8, a_tol=0.0, max_iter=4000, check_every=8)
        pid = alone.add(p["op64"], p["op32"], p["d32"], p["x"], p["b"])
        alone.solve()
        np.testing.assert_array_equal(p["x"].export_field(), mixed[i])
        alone.remove(pid)
        alone.close()


def test_served_configuration_stream_slice(hip):
    import ctypes
    from lfric_core_amd import serving
    from lfric_core_amd.batch import Batch
    reqs = serving.poisson_stream(200.0, 60.0, seed=0)[:12]
    assert [r.rid for r in reqs] == list(range(12))
    assert {r.n for r in reqs} <= {24, 48, 96} and {r.tol for r in reqs} == {1e-6, 1e-8}
    free, total = ctypes.c_longlong(), ctypes.c_longlong()
    hip.lib().lfric_device_memory(ctypes.byref(free), ctypes.byref(total))
    budget = min(288 * 10**9, free.value) - 8 * 10**9
    srv = serving.SolveServer(budget, starve_s=2.0, check_every=8)
    keep = {}
    orig_live = serving._Live

    class Recording(orig_live):
        def __init__(self, req):
            super().__init__(req)
            keep[req.rid] = self

    serving._Live = Recording
    try:
        rep = srv.serve(reqs, wall_limit_s=60.0)
    finally:
        serving._Live = orig_live
    print(f"served {rep.completed} requests, p50 {rep.p50_ms:.0f} ms, p99 {rep.p99_ms:.0f} ms", flush=True)
    assert rep.completed == len(reqs) and rep.rejected == 0 and rep.peak_bytes <= budget
    for r in reqs:
        assert r.status == 0 and r.true_rel_res <= r.tol, (r.rid, r.status, r.true_rel_res, r.tol)
    # three sampled requests: the first, a C96 order-0 and a C96 order-1 one where present
    sample = [reqs[0].rid]
    for order in (0, 1):
        c = [r.rid for r in reqs if r.n == 96 and r.order == order]
        if c:
            sample.append(c[0])
    assert len(sample) == 3
    for rid in sample:
        r = reqs[rid]
        mixed = keep[rid].x.export_field().copy()
        del keep[rid]
        alone = serving._Live(r)
        bt = Batch(r_tol=r.tol, a_tol=0.0, check_every=8, inner_r_tol=1e-4, max_iter=4000)
        pid = bt.add(alone.op64, alone.op32, alone.diag32, alone.x, alone.b)
        bt.solve()
        np.testing.assert_array_equal(alone.x.export_field(), mixed)
        bt.remove(pid)
        bt.close()
        del alone
    keep.clear()
    srv.close()
