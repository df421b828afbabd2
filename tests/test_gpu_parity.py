"""GPU parity: the HIP path (through the C ABI) against the oracle on the same seeded inputs.

Bars: bit-exact for integer work (dofmaps, colouring) and for the counter-based generator;
FP64 LMA apply within 1e-13 relative of the oracle (element-wise, relative to |A||x|); FP32
apply within 2e-6 relative to |A||x| (FP32 accumulation over <= 2*ndf terms); CG solutions
within the stopping tolerance times the operator's conditioning (stated per test)."""
import json

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _mesh_pair(hip, orc, n, nl, space=0, order=0):
    fs = hip.FunctionSpace.cubedsphere(n, nl, space, order)
    m = orc.Mesh(n, nl, space, order)
    return fs, m


@pytest.mark.parametrize("n,nl,space,order", [(4, 3, 0, 0), (12, 20, 0, 0), (4, 5, 0, 1), (12, 5, 2, 0),
                                              (12, 5, 3, 0), (96, 2, 0, 0), (3, 2, 0, 0), (5, 3, 2, 0),
                                              (7, 2, 0, 1)])
def test_dofmap_and_colouring_bit_exact(hip, orc, n, nl, space, order):
    fs, m = _mesh_pair(hip, orc, n, nl, space, order)
    assert (fs.ncells, fs.ndf, fs.undf) == (m.ncells, m.ndf, m.undf)
    np.testing.assert_array_equal(fs.dofmap(), m.dofmap)
    npc, cic = fs.colouring()
    np.testing.assert_array_equal(npc, m.ncells_per_colour)
    np.testing.assert_array_equal(cic, m.cells_in_colour)


@pytest.mark.parametrize("ndf_order,scaled", [(0, False), (0, True), (1, True)])
def test_generator_bit_exact(hip, orc, ndf_order, scaled):
    fs, m = _mesh_pair(hip, orc, 4, 3, 0, ndf_order)
    op = hip.LmaOperator(fs).generate_spd(11, scaled=scaled)
    A = orc.gen_spd(m.ndf, m.ncell_3d, 11, scaled=scaled)
    np.testing.assert_array_equal(op.download(), A)


def test_upload_download_round_trip(hip, orc):
    fs, m = _mesh_pair(hip, orc, 4, 3)
    A = np.random.default_rng(0).standard_normal((m.ncell_3d, 8, 8))
    op = hip.LmaOperator(fs).upload(A)
    np.testing.assert_array_equal(op.download(), A)
    # precision changes move entries between the FP64 and FP32 chunk layouts
    op32 = op.convert(hip.R32)
    np.testing.assert_array_equal(op32.download(), A.astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(op32.convert(hip.R64).download(), A.astype(np.float32).astype(np.float64))
    # order 1 (27x27 entries, not a whole number of chunks per row)
    fs1, m1 = _mesh_pair(hip, orc, 4, 2, 0, 1)
    A1 = np.random.default_rng(1).standard_normal((m1.ncell_3d, 27, 27))
    op1 = hip.LmaOperator(fs1).upload(A1)
    np.testing.assert_array_equal(op1.download(), A1)
    np.testing.assert_array_equal(op1.convert(hip.R32).download(), A1.astype(np.float32).astype(np.float64))


def _bound(orc, m, A, x, rel):
    S = orc.assembled_csr(m, A)
    return rel * (abs(S) @ np.abs(x))


@pytest.mark.parametrize("method", ["apply", "apply_gather"])
@pytest.mark.parametrize("n,nl,space,order", [(4, 3, 0, 0), (12, 20, 0, 0), (4, 5, 0, 1), (12, 7, 0, 1),
                                              (6, 4, 2, 0), (5, 3, 0, 0)])
def test_lma_apply_fp64_parity(hip, orc, n, nl, space, order, method):
    fs, m = _mesh_pair(hip, orc, n, nl, space, order)
    rng = np.random.default_rng(n + nl)
    A = rng.standard_normal((m.ncell_3d, m.ndf, m.ndf))
    x = rng.standard_normal(m.undf)
    op = hip.LmaOperator(fs).upload(A)
    xv = hip.FieldVector(fs).import_field(x)
    yv = hip.FieldVector(fs)
    yv.set_scalar(12345.0)  # the apply must overwrite, not accumulate
    getattr(op, method)(xv, yv)
    y = yv.export_field()
    yo = orc.lma_apply(m, A, x)
    assert np.all(np.abs(y - yo) <= _bound(orc, m, A, x, 1e-13))
    # run-to-run bitwise determinism (no atomics)
    yv.set_scalar(-1.0)
    getattr(op, method)(xv, yv)
    np.testing.assert_array_equal(yv.export_field(), y)


@pytest.mark.parametrize("method", ["apply", "apply_gather"])
@pytest.mark.parametrize("order", [0, 1])
def test_lma_apply_fp32_parity(hip, orc, order, method):
    fs, m = _mesh_pair(hip, orc, 12, 10, 0, order)
    A = orc.gen_spd(m.ndf, m.ncell_3d, 3, scaled=True)
    x = orc.gen_uniform(m.undf, 0, 3)
    op64 = hip.LmaOperator(fs).generate_spd(3, scaled=True)
    op32 = op64.convert(hip.R32)
    xv = hip.FieldVector(fs, hip.R32).import_field(x)
    yv = hip.FieldVector(fs, hip.R32)
    getattr(op32, method)(xv, yv)
    y = yv.export_field()
    yo = orc.lma_apply(m, A.astype(np.float32), x.astype(np.float32), r32=True).astype(np.float64)
    assert np.all(np.abs(y - yo) <= _bound(orc, m, A, x, 2e-6))


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n,nl,order,nrhs", [(4, 3, 0, 4), (12, 7, 0, 2), (4, 3, 1, 4), (5, 4, 0, 8)])
def test_lma_apply_multi_rhs(hip, orc, n, nl, order, nrhs, prec):
    """Column r of the multi-RHS apply equals the one-RHS apply of column r bit for bit (same
    variant), and the oracle within the FP64 / FP32 bars."""
    fs, m = _mesh_pair(hip, orc, n, nl, 0, order)
    op64 = hip.LmaOperator(fs).generate_spd(2, scaled=True)
    op = op64 if prec == 0 else op64.convert(hip.R32)
    P = hip.R64 if prec == 0 else hip.R32
    X = np.random.default_rng(nrhs).uniform(-1, 1, (m.undf, nrhs))
    if prec:
        X = X.astype(np.float32).astype(np.float64)
    xv = hip.FieldVector(fs, P, nrhs).import_field(X.ravel())
    yv = hip.FieldVector(fs, P, nrhs)
    A = orc.gen_spd(m.ndf, m.ncell_3d, 2, scaled=True)
    for variant, single in ((1, "apply_gather"), (2, "apply")):
        yv.set_scalar(3.0)
        op.apply_mrhs(xv, yv, variant)
        Y = yv.export_field().reshape(m.undf, nrhs)
        for r in range(nrhs):
            x1 = hip.FieldVector(fs, P).import_field(X[:, r])
            y1 = hip.FieldVector(fs, P)
            getattr(op, single)(x1, y1)
            np.testing.assert_array_equal(Y[:, r], y1.export_field())
            yo = orc.lma_apply(m, A, X[:, r])
            assert np.all(np.abs(Y[:, r] - yo) <= _bound(orc, m, A, X[:, r], 1e-13 if prec == 0 else 2e-6))


@pytest.mark.parametrize("n,nl,nrhs", [(4, 3, 1), (4, 3, 4), (12, 7, 2), (5, 4, 8), (6, 2, 4)])
def test_multi_rhs_mfma_is_bit_identical(hip, orc, n, nl, nrhs):
    """The matrix-core multi-RHS apply (chains of v_mfma_f32_4x4x1f32) adds the same products in
    the same order as the colour-ordered FMA kernel, so the two agree bit for bit, and both lie
    within the FP32 bar of the oracle.  Operators it does not cover are refused."""
    fs, m = _mesh_pair(hip, orc, n, nl)
    op = hip.LmaOperator(fs).generate_spd(3, scaled=True).convert(hip.R32)
    X = np.random.default_rng(n * 10 + nrhs).uniform(-1, 1, (m.undf, nrhs)).astype(np.float32).astype(np.float64)
    xv = hip.FieldVector(fs, hip.R32, nrhs).import_field(X.ravel())
    y2, y3 = hip.FieldVector(fs, hip.R32, nrhs), hip.FieldVector(fs, hip.R32, nrhs)
    y2.set_scalar(1.0)
    y3.set_scalar(-2.0)
    op.apply_mrhs(xv, y2, 2)
    op.apply_mrhs(xv, y3, 3)
    np.testing.assert_array_equal(y3.export_raw(), y2.export_raw())
    A = orc.gen_spd(8, m.ncell_3d, 3, scaled=True)
    Y = y3.export_field().reshape(m.undf, nrhs)
    for r in range(nrhs):
        yo = orc.lma_apply(m, A.astype(np.float32), X[:, r].astype(np.float32), r32=True).astype(np.float64)
        assert np.all(np.abs(Y[:, r] - yo) <= _bound(orc, m, A, X[:, r], 2e-6))
    op64 = hip.LmaOperator(fs).generate_spd(3)
    with pytest.raises(hip.LfricError):
        op64.apply_mrhs(hip.FieldVector(fs, hip.R64, nrhs), hip.FieldVector(fs, hip.R64, nrhs), 3)


@pytest.mark.parametrize("prec", [0, 1])
def test_row_gathered_copy_is_bit_identical(hip, orc, prec):
    """The gather apply over the row-gathered copy adds the same products in the same order as
    over the cell layout, for one and for several right-hand sides; CG with the fused p.(A p)
    then takes the same iterations."""
    fs, m = _mesh_pair(hip, orc, 12, 9)
    op64 = hip.LmaOperator(fs).generate_spd(4)
    op = op64 if prec == 0 else op64.convert(hip.R32)
    P = hip.R64 if prec == 0 else hip.R32
    xv = hip.FieldVector(fs, P).fill_uniform(2, 3)
    y0, y1 = hip.FieldVector(fs, P), hip.FieldVector(fs, P)
    op.apply_gather(xv, y0)
    before = y0.export_field()
    xm = {nr: hip.FieldVector(fs, P, nr).fill_uniform(5 + nr, 3) for nr in (2, 4, 8)}
    ym = {nr: hip.FieldVector(fs, P, nr) for nr in (2, 4, 8)}
    mbefore = {}
    for nr in (2, 4, 8):
        op.apply_mrhs(xm[nr], ym[nr], 1)
        mbefore[nr] = ym[nr].export_field()
    op.prepare_rows()
    for nr in (2, 4, 8):
        ym[nr].set_scalar(7.0)
        op.apply_mrhs(xm[nr], ym[nr], 1)
        np.testing.assert_array_equal(ym[nr].export_field(), mbefore[nr])
    for w in (0, 4):
        op.set_window(w)
        y1.set_scalar(5.0)
        op.apply_gather(xv, y1)
        np.testing.assert_array_equal(y1.export_field(), before)
    A = orc.gen_spd(8, m.ncell_3d, 4)
    yo = orc.lma_apply(m, A, xv.export_field())
    assert np.all(np.abs(before - yo) <= _bound(orc, m, A, xv.export_field(), 1e-13 if prec == 0 else 2e-6))
    if prec == 0:
        b = hip.FieldVector(fs).fill_uniform(0, 3)
        prec_ = hip.DiagonalPreconditioner.from_operator(op)
        x = hip.FieldVector(fs)
        x.set_scalar(0.0)
        res = hip.ConjugateGradient(op, prec_, 1e-8, 1e-16, 2000, check_every=4).apply(x, b)
        d = orc.mm_diagonal(m, A)
        xo, ro, _ = orc.cg(m, A, b.export_field(), diag=d, max_iter=2000)
        assert res.iterations == ro.iterations
        assert np.linalg.norm(x.export_field() - xo) <= 1e-8 * np.linalg.norm(xo)


def test_lma_apply_known_answer(hip):
    k = json.loads((GOLDEN / "kat.json").read_text())["matrix_vector"]
    colours = dict(ncells_per_colour=[1], cells_in_colour=[[1]])
    to = hip.FunctionSpace.from_dofmap(1, 4, [k["map1"]], **colours)
    fr = hip.FunctionSpace.from_dofmap(1, 2, [k["map2"]], **colours)
    op = hip.LmaOperator(to, fr).upload(np.array(k["d"])[None])
    x = hip.FieldVector(fr).import_field(k["x"])
    y = hip.FieldVector(to)
    for method in ("apply", "apply_gather"):
        y.set_scalar(7.0)
        getattr(op, method)(x, y)
        np.testing.assert_allclose(y.export_field(), k["answer"], rtol=0, atol=k["tol"])


def test_mm_diagonal_parity(hip, orc):
    fs, m = _mesh_pair(hip, orc, 12, 20)
    op = hip.LmaOperator(fs).generate_spd(0)
    d = op.mm_diagonal(hip.FieldVector(fs)).export_field()
    np.testing.assert_allclose(d, orc.mm_diagonal(m, orc.gen_spd(8, m.ncell_3d, 0)), rtol=1e-14)


def test_vector_ops_against_numpy(hip, orc):
    fs = hip.FunctionSpace.cubedsphere(12, 20)
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(fs.undf), rng.standard_normal(fs.undf)
    x, y = hip.FieldVector(fs).import_field(a), hip.FieldVector(fs).import_field(b)
    assert abs(x.dot(y) - a @ b) <= 1e-13 * (np.abs(a) @ np.abs(b))
    assert abs(x.norm() - np.linalg.norm(a)) <= 1e-13 * np.linalg.norm(a)
    z = x.duplicate()
    z.axpby(2.0, x, -3.0, y)
    np.testing.assert_allclose(z.export_field(), 2 * a - 3 * b, rtol=1e-15, atol=1e-15)
    z.aypx(0.5, x)
    np.testing.assert_allclose(z.export_field(), 0.5 * (2 * a - 3 * b) + a, rtol=1e-15, atol=1e-15)
    z.axpy(-1.5, y)
    z.scale(2.0)
    np.testing.assert_allclose(z.export_field(), 2 * (0.5 * (2 * a - 3 * b) + a - 1.5 * b), rtol=1e-14, atol=1e-14)
    z.copy(x)
    np.testing.assert_array_equal(z.export_field(), a)
    z.divide(x, y)
    np.testing.assert_array_equal(z.export_field(), a / b)
    # FP32 storage with FP64-accumulated dots (r_solver_field_vector_type)
    x32 = hip.FieldVector(fs, hip.R32).import_field(a)
    a32 = a.astype(np.float32).astype(np.float64)
    assert abs(x32.dot(x32) - a32 @ a32) <= 1e-12 * (a32 @ a32)


def _config0(hip, orc, n=12, nl=20, scaled=False):
    fs, m = _mesh_pair(hip, orc, n, nl)
    op = hip.LmaOperator(fs).generate_spd(0, scaled=scaled)
    b = hip.FieldVector(fs).fill_uniform(0, 3)
    return fs, m, op, b


@pytest.mark.parametrize("form", [0, 1])
def test_cg_fp64_matches_oracle(hip, orc, form):
    fs, m, op, b = _config0(hip, orc)
    prec = hip.DiagonalPreconditioner.from_operator(op)
    x = hip.FieldVector(fs)
    x.set_scalar(0.0)
    cg = hip.ConjugateGradient(op, prec, 1e-8, 1e-16, 500, check_every=4)
    hip.set_apply_form(form)  # 1: gather apply with the fused p.(A p)
    try:
        res = cg.apply(x, b)
    finally:
        hip.set_apply_form(1)  # the library default
    A = orc.gen_spd(8, m.ncell_3d, 0)
    bo = orc.gen_uniform(m.undf, 0, 3)
    np.testing.assert_array_equal(b.export_field(), bo)
    xo, reso, hist = orc.cg(m, A, bo, diag=orc.mm_diagonal(m, A))
    assert res.status == 0 and reso.converged == 1
    assert res.iterations == reso.iterations
    np.testing.assert_allclose(res.r_nrm / res.r_nrm_0, reso.r_nrm / reso.r_nrm_0, rtol=1e-6)
    xg = x.export_field()
    assert np.linalg.norm(xg - xo) <= 1e-10 * np.linalg.norm(xo)
    S = orc.assembled_csr(m, A)
    assert np.linalg.norm(bo - S @ xg) / np.linalg.norm(bo) <= 1.01e-8


@pytest.mark.parametrize("rows", [False, True])
def test_cg_iterates_do_not_depend_on_the_xcd_window(hip, orc, rows):
    """The fused p.(A p) partials are stored by the logical block the gather kernels work on, so
    the order of CG's dot, and with it every iterate, is the same for every block-order window."""
    fs, m, op, b = _config0(hip, orc, 12, 20, scaled=True)
    op32 = op.convert(hip.R32)
    if rows:
        op32.prepare_rows()
    prec = hip.DiagonalPreconditioner.from_operator(op32)
    b32 = hip.FieldVector(fs, hip.R32).import_field(b.export_field())
    out = []
    for window in (0, 8, 2):
        op32.set_window(window)
        x = hip.FieldVector(fs, hip.R32)
        x.set_scalar(0.0)
        res = hip.ConjugateGradient(op32, prec, 1e-6, 0.0, 500).apply(x, b32)
        out.append((res.iterations, res.r_nrm, x.export_raw()))
    for it, nrm, xw in out[1:]:
        assert it == out[0][0] and nrm == out[0][1]
        assert np.array_equal(xw.view(np.uint32), out[0][2].view(np.uint32))


@pytest.mark.parametrize("form", [0, 1])
def test_mixed_precision_refinement_reaches_fp64_accuracy(hip, orc, form):
    fs, m, op, b = _config0(hip, orc, 12, 20, scaled=True)
    op32 = op.convert(hip.R32)
    prec32 = hip.DiagonalPreconditioner.from_operator(op32)
    x = hip.FieldVector(fs)
    solver = hip.MixedPrecisionCG(op, op32, prec32, 1e-8, 0.0, 5000, inner_r_tol=1e-4)
    hip.set_apply_form(form)
    try:
        res = solver.apply(x, b)
    finally:
        hip.set_apply_form(1)  # the library default
    assert res.status == 0 and res.true_rel_res <= 1e-8 and res.outer_iterations >= 2
    A = orc.gen_spd(8, m.ncell_3d, 0, scaled=True)
    bo = orc.gen_uniform(m.undf, 0, 3)
    S = orc.assembled_csr(m, A)
    xg = x.export_field()
    assert np.linalg.norm(bo - S @ xg) / np.linalg.norm(bo) <= 1.01e-8
    # against the FP64 oracle solve driven to 1e-12.  The refined solution's error is about 1e-8
    # here; an inner FP32 error left uncorrected would sit near FP32 precision times the
    # condition number, orders of magnitude above this bar
    xo, reso, _ = orc.cg(m, A, bo, diag=orc.mm_diagonal(m, A), r_tol=1e-12, max_iter=20000)
    err = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    print(f"mixed-precision solution error vs the FP64 oracle: {err:.3e}")
    assert err <= 1e-6, err


def test_solver_status_codes(hip, orc):
    fs, m, op, b = _config0(hip, orc, 4, 3)
    prec = hip.DiagonalPreconditioner.from_operator(op)
    x = hip.FieldVector(fs)
    # zero right-hand side: converged in 0 iterations (||r0|| < EPS)
    zero = hip.FieldVector(fs)
    res = hip.ConjugateGradient(op, prec, 1e-8, 0.0, 50).apply(x, zero)
    assert res.iterations == 0 and res.status == 0
    # max_iter reached with fail_on_non_converged raises NotConverged
    with pytest.raises(hip.NotConverged):
        hip.ConjugateGradient(op, prec, 1e-14, 0.0, 1, fail_on_non_converged=True).apply(x, b)
    # NaN in b is reported as NaN, not as convergence
    bn = np.full(fs.undf, 1.0)
    bn[5] = np.nan
    with pytest.raises(hip.LfricError) as e:
        hip.ConjugateGradient(op, prec, 1e-8, 0.0, 10).apply(hip.FieldVector(fs), hip.FieldVector(fs).import_field(bn))
    assert e.value.code == 2


@pytest.mark.parametrize("space,order,prec", [(3, 0, 0), (2, 0, 0), (0, 0, 0), (0, 1, 0), (3, 0, 1),
                                              (2, 0, 1), (0, 0, 1), (0, 1, 1)])
def test_invert_local_bitwise_equals_oracle(hip, orc, space, order, prec):
    """Per-cell dense inverse (invert_local_operator_code with matrix_invert_lu): the device
    repeats the reference's operation order without contraction, so every entry equals the
    oracle's (itself bit-equal to the reference's compiled matrix_invert_lu) for FP64 and FP32;
    spaces with and without top partners, register (<= 8 dofs) and workspace (27 dofs) forms."""
    fs, m = _mesh_pair(hip, orc, 4, 3, space, order)
    rng = np.random.default_rng(100 * space + 10 * order + prec)
    A = rng.uniform(-1, 1, (m.ncell_3d, m.ndf, m.ndf)) + m.ndf * np.eye(m.ndf)
    op = hip.LmaOperator(fs).upload(A)
    if prec:
        op = op.convert(hip.R32)
        A = A.astype(np.float32).astype(np.float64)
    inv = op.invert_local().download()
    np.testing.assert_array_equal(inv, orc.invert_local(A, r32=bool(prec)).astype(np.float64))


def test_invert_local_known_answer_and_refusals(hip, orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["invert_local_operator"]
    fs, m = _mesh_pair(hip, orc, 3, k["nlayers"], 3, 0)
    op = hip.LmaOperator(fs).upload(np.full((m.ncell_3d, 1, 1), k["value"]))
    inv = op.invert_local().download()
    assert np.all(np.abs(inv - k["answer"]) <= k["tol"])
    w0 = hip.FunctionSpace.cubedsphere(3, k["nlayers"], 0, 0)
    with pytest.raises(hip.LfricError):  # an operator between two spaces has no local inverse
        hip.LmaOperator(fs, w0).invert_local()


def test_row_copy_follows_operator_changes(hip, orc):
    """An operator rewritten after its row-gathered copy was built applies its new entries."""
    fs, m = _mesh_pair(hip, orc, 6, 4)
    op = hip.LmaOperator(fs).generate_spd(1).convert(hip.R32).prepare_rows()
    fresh = hip.LmaOperator(fs).generate_spd(2).convert(hip.R32)
    x = hip.FieldVector(fs, hip.R32).fill_uniform(3)
    y0, y1 = hip.FieldVector(fs, hip.R32), hip.FieldVector(fs, hip.R32)
    hip.LmaOperator(fs).generate_spd(2).convert(hip.R32)  # unrelated operator: no effect
    op.upload(fresh.download())
    op.apply_gather(x, y0)
    fresh.apply_gather(x, y1)
    np.testing.assert_array_equal(y0.export_raw(), y1.export_raw())


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n,nl,seg", [(12, 9, 4), (12, 9, 8), (5, 4, 3), (6, 7, 2), (6, 7, 1), (4, 3, 64),
                                      (12, 9, 0), (5, 4, 0), (6, 70, 0), (4, 2, 0)])
def test_symmetric_packed_apply_equals_colour_apply(hip, orc, n, nl, seg, prec):
    """The symmetric-packed apply (upper triangles, per-cell outputs, colour-ordered gather) forms
    the colour-ordered scatter's sums in its order: equal bit for bit, for any segment length;
    CG over it takes the oracle's iteration count.  Non-symmetric operators are refused, and an
    operator rewritten with non-symmetric entries drops the copy."""
    fs, m = _mesh_pair(hip, orc, n, nl)
    op64 = hip.LmaOperator(fs).generate_spd(6, scaled=True)
    op = op64 if prec == 0 else op64.convert(hip.R32)
    P = hip.R64 if prec == 0 else hip.R32
    x = hip.FieldVector(fs, P).fill_uniform(4, 3)
    yc, ys = hip.FieldVector(fs, P), hip.FieldVector(fs, P)
    op.apply(x, yc)
    op.prepare_symmetric(seg)
    ys.set_scalar(9.0)
    op.apply_gather(x, ys)
    np.testing.assert_array_equal(ys.export_raw(), yc.export_raw())
    if prec == 0 and seg in (4, 0) and n == 12:
        b = hip.FieldVector(fs).fill_uniform(0, 3)
        xs = hip.FieldVector(fs)
        xs.set_scalar(0.0)
        res = hip.ConjugateGradient(op, hip.DiagonalPreconditioner.from_operator(op), 1e-8, 1e-16, 2000,
                                    check_every=4).apply(xs, b)
        A = orc.gen_spd(8, m.ncell_3d, 6, scaled=True)
        xo, ro, _ = orc.cg(m, A, b.export_field(), diag=orc.mm_diagonal(m, A), max_iter=2000)
        assert res.iterations == ro.iterations
        assert np.linalg.norm(xs.export_field() - xo) <= 1e-8 * np.linalg.norm(xo)
    nonsym = hip.LmaOperator(fs, precision=P).generate_uniform(1, 5, -1.0, 1.0)
    with pytest.raises(hip.LfricError):
        nonsym.prepare_symmetric(seg)
    # rewriting with non-symmetric entries drops the copy: the gather apply is the plain one again
    A_ns = nonsym.download()
    op.upload(A_ns)
    y1, y2 = hip.FieldVector(fs, P), hip.FieldVector(fs, P)
    op.apply_gather(x, y1)
    nonsym.apply_gather(x, y2)
    np.testing.assert_array_equal(y1.export_raw(), y2.export_raw())


def test_apply_and_solve_on_a_caller_stream(hip, orc):
    """lfric_lma_apply_stream and lfric_set_stream run the same kernels on a stream the caller
    created (default priority): the results equal those on the library's stream bit for bit."""
    import ctypes
    from lfric_core_amd import _lib
    L = _lib.lib()
    hiprt = ctypes.CDLL("libamdhip64.so")
    stream = ctypes.c_void_p()
    assert hiprt.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        fs, m, op, b = _config0(hip, orc, 8, 6)
        y0, y1 = hip.FieldVector(fs), hip.FieldVector(fs)
        op.apply(b, y0)
        assert L.lfric_lma_apply_stream(op._h, b._h, y1._h, stream) == 0
        assert hiprt.hipStreamSynchronize(stream) == 0
        np.testing.assert_array_equal(y1.export_field(), y0.export_field())
        prec = hip.DiagonalPreconditioner.from_operator(op)
        xs = []
        for s in (None, stream):
            assert L.lfric_set_stream(s) == 0
            x = hip.FieldVector(fs)
            x.set_scalar(0.0)
            res = hip.ConjugateGradient(op, prec, 1e-8, 0.0, 500).apply(x, b)
            xs.append((res.iterations, x.export_field()))
        assert L.lfric_set_stream(None) == 0
        assert xs[0][0] == xs[1][0]
        np.testing.assert_array_equal(xs[0][1], xs[1][1])
    finally:
        L.lfric_set_stream(None)
        hiprt.hipStreamDestroy(stream)
