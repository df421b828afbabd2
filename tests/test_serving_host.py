"""Host logic of the served stream (CPU only): the seeded Poisson request stream, the request
footprints and the admission rule."""
import numpy as np

from lfric_core_amd import serving


def _req(rid, t, fp):
    r = serving.Request(rid, t, 24, 30, 0, 1, 1e-8, rid)
    r.footprint = fp
    return r


def test_poisson_stream_is_seeded_and_has_the_configured_mix():
    a = serving.poisson_stream(200.0, 60.0, seed=0)
    b = serving.poisson_stream(200.0, 60.0, seed=0)
    assert [(r.t_arrive, r.n, r.nl, r.order, r.nrhs, r.tol, r.seed) for r in a] == \
           [(r.t_arrive, r.n, r.nl, r.order, r.nrhs, r.tol, r.seed) for r in b]
    assert serving.poisson_stream(200.0, 60.0, seed=1)[0].t_arrive != a[0].t_arrive
    n = len(a)
    assert abs(n - 12000) < 5 * np.sqrt(12000)  # Poisson count, mean rate x window
    t = np.array([r.t_arrive for r in a])
    assert np.all(np.diff(t) > 0) and t[-1] < 60.0
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 200.0) < 5 * gaps.std() / np.sqrt(len(gaps))
    meshes = np.array([r.n for r in a])
    for m, w in zip((24, 48, 96, 192), (0.4, 0.3, 0.2, 0.1)):
        assert abs((meshes == m).mean() - w) < 0.02
    assert {r.nl for r in a} == {30, 70} and {r.order for r in a} == {0, 1}
    assert {r.nrhs for r in a} == set(range(1, 9)) and {r.tol for r in a} == {1e-6, 1e-8}


def test_footprint_counts_follow_the_spaces(orc):
    for n, nl, order in ((3, 2, 0), (4, 3, 1), (5, 2, 1)):
        r = serving.Request(0, 0.0, n, nl, order, 3, 1e-8, 0)
        assert serving.footprint_undf(r) == orc.Mesh(n, nl, orc.W0, order).undf
    assert [serving.padded_rhs(k) for k in range(1, 9)] == [1, 2, 4, 4, 8, 8, 8, 8]
    # an order-1 C192x70 request holds ~137 GB of FP64 + FP32 operators and ~27 GB of the FP32
    # operator's symmetric-packed copy and row sums
    assert 1.65e11 < serving.footprint_bytes(192, 70, 1, 1) < 1.80e11
    assert serving.footprint_bytes(24, 30, 0, 8) > serving.footprint_bytes(24, 30, 0, 1)


def test_admission_backfills_until_the_head_starves():
    big, small1, small2 = _req(0, 0.0, 80), _req(1, 0.1, 10), _req(2, 0.2, 10)
    waiting = [big, small1, small2]
    # 50 of 100 bytes in use: the head does not fit, later requests go ahead of it
    assert serving.admissions(waiting, 50, 100, now=1.0, starve_s=2.0) == [small1, small2]
    # the head has waited more than starve_s: nothing is admitted behind it
    assert serving.admissions(waiting, 50, 100, now=3.0, starve_s=2.0) == []
    # once it fits it goes first, and the budget is respected for the rest
    assert serving.admissions(waiting, 10, 100, now=3.0, starve_s=2.0) == [big, small1]
    assert serving.admissions([], 0, 100, now=0.0, starve_s=2.0) == []
