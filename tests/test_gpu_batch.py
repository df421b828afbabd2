"""GPU tests of the batched solver (lfric_batch_*): uneven problems (meshes, levels, element
orders, right-hand sides) solved together in FP32 CG inside FP64 refinement.

Bars: every RHS reaches a true FP64 relative residual <= 1e-8 against the oracle's assembled
matrix; each problem's solution is bit-identical when solved alone, in the mix, with or without
the HIP graph, after autotuning, and when it joins a running batch later (continuous
batching)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (edge cells, levels, element order, nrhs, seed)
MIX = [(4, 3, 0, 4, 1), (5, 4, 1, 2, 2), (6, 3, 0, 1, 3), (4, 2, 1, 8, 4), (8, 5, 0, 4, 5)]


def _make(hip, n, nl, order, nrhs, seed):
    fs = hip.FunctionSpace.cubedsphere(n, nl, hip.W0, order)
    op64 = hip.LmaOperator(fs).generate_spd(seed, scaled=True)
    op32 = op64.convert(hip.R32)
    diag32 = op32.mm_diagonal(hip.FieldVector(fs, hip.R32))
    B = np.random.default_rng(seed).uniform(-1, 1, (fs.undf, nrhs))
    b = hip.FieldVector(fs, hip.R64, nrhs).import_field(B.ravel())
    x = hip.FieldVector(fs, hip.R64, nrhs)
    x.set_scalar(0.0)
    return dict(fs=fs, op64=op64, op32=op32, diag32=diag32, b=b, x=x, B=B, key=(n, nl, order, nrhs, seed))


def _solve(hip, probs, graph=True, tune=False, late=(), colour=True):
    from lfric_core_amd.batch import Batch
    bt = Batch(r_tol=1e-8, check_every=4)
    bt.config(use_graph=1 if graph else 0, colour=colour)
    ids = {}
    for i, p in enumerate(probs):
        if i in late:
            continue
        p["x"].set_scalar(0.0)
        ids[i] = bt.add(p["op64"], p["op32"], p["diag32"], p["x"], p["b"])
    if tune:
        bt.run(1)
        assert bt.tune() > 0
    if late:
        bt.run(2)
        for i in late:
            probs[i]["x"].set_scalar(0.0)
            ids[i] = bt.add(probs[i]["op64"], probs[i]["op32"], probs[i]["diag32"], probs[i]["x"], probs[i]["b"])
    bt.solve()
    out = {}
    for i, pid in ids.items():
        assert bt.finished(pid)
        out[i] = (probs[i]["x"].export_field().copy(), bt.result(pid))
        bt.remove(pid)
    bt.close()
    return out


def test_batch_mix_converges_and_matches_alone(hip, orc):
    probs = [_make(hip, *k) for k in MIX]
    mixed = _solve(hip, probs)
    for i, p in enumerate(probs):
        n, nl, order, nrhs, seed = p["key"]
        m = orc.Mesh(n, nl, orc.W0, order)
        S = orc.assembled_csr(m, orc.gen_spd(m.ndf, m.ncell_3d, seed, scaled=True))
        X = mixed[i][0].reshape(m.undf, nrhs)
        for r, res in enumerate(mixed[i][1]):
            assert res.status == 0, (p["key"], r, res)
            rel = np.linalg.norm(p["B"][:, r] - S @ X[:, r]) / np.linalg.norm(p["B"][:, r])
            assert rel <= 1e-8, (p["key"], r, rel)
            assert res.true_rel_res <= 1e-8
        alone = _solve(hip, [p])[0]
        np.testing.assert_array_equal(alone[0], mixed[i][0])
        assert [r.iterations for r in alone[1]] == [r.iterations for r in mixed[i][1]]


def test_batch_graph_tuning_and_late_joiners_are_bit_identical(hip, orc):
    probs = [_make(hip, *k) for k in MIX[:4]]
    ref = _solve(hip, probs, graph=False)
    for kw in (dict(graph=True), dict(graph=True, tune=True), dict(graph=True, late=(1, 3))):
        got = _solve(hip, probs, **kw)
        for i in ref:
            np.testing.assert_array_equal(got[i][0], ref[i][0])


def test_batch_gather_form_converges_and_agrees(hip, orc):
    """The batch with every problem in gather form (instead of the colour-ordered matrix-core /
    register-staged forms) reaches the same FP64 bar; both solutions agree to the conditioning
    of the systems (unscaled element matrices, condition number of a few hundred)."""
    probs = []
    for k in MIX:
        p = _make(hip, *k)
        p["op64"].generate_spd(k[-1], scaled=False)
        p["op32"] = p["op64"].convert(hip.R32)
        p["diag32"] = p["op32"].mm_diagonal(hip.FieldVector(p["fs"], hip.R32))
        probs.append(p)
    col = _solve(hip, probs)
    gat = _solve(hip, probs, colour=False)
    for i, p in enumerate(probs):
        for res in col[i][1] + gat[i][1]:
            assert res.status == 0 and res.true_rel_res <= 1e-8, (p["key"], res)
        d = np.linalg.norm(col[i][0] - gat[i][0]) / np.linalg.norm(gat[i][0])
        assert d <= 1e-5, (p["key"], d)


def test_batch_copy_forms_equal_colour_and_gather_forms(hip, orc):
    """The batch's default forms over the operators' copies: the symmetric-packed order-1 apply
    equals the colour-ordered register-staged one bit for bit, and the gather over the
    row-gathered copy equals the gather form bit for bit, so each problem's solution and
    iteration counts are those of the batch run in that form."""
    from lfric_core_amd.batch import prepare_batch_copies
    probs = [_make(hip, *k) for k in MIX + [(6, 5, 1, 4, 6), (3, 3, 1, 1, 7)]]
    for p in probs:
        prepare_batch_copies(p["op32"])

    def run(apply):
        from lfric_core_amd.batch import Batch
        bt = Batch(r_tol=1e-8, check_every=4)
        bt.config(apply=apply)
        ids = {}
        for i, p in enumerate(probs):
            p["x"].set_scalar(0.0)
            ids[i] = bt.add(p["op64"], p["op32"], p["diag32"], p["x"], p["b"])
        bt.solve()
        out = {i: (probs[i]["x"].export_field().copy(), bt.result(pid)) for i, pid in ids.items()}
        bt.close()
        return out

    cop, col, gat = run("copies"), run("colour"), run("gather")
    for i, p in enumerate(probs):
        order = p["key"][2]
        for res in cop[i][1]:
            assert res.status == 0 and res.true_rel_res <= 1e-8, (p["key"], res)
        same = col if order == 1 else gat
        np.testing.assert_array_equal(cop[i][0], same[i][0])
        assert [r.iterations for r in cop[i][1]] == [r.iterations for r in same[i][1]]
