"""GPU parity of the mass-matrix operator's boundary conditions and the dg_inc apply.

enforce_bc (sci_enforce_bc_kernel_mod.F90:63-91) only multiplies by 0 or 1, so the device result
must equal the oracle's bit for bit; dg_inc (dg_inc_matrix_vector_kernel_mod.F90:69-142) is
compared within 1e-13 (FP64) / 2e-6 (FP32) of |A||x| + |y0| norm-wise; the W2 mass-matrix CG
with boundary conditions takes the oracle's iteration count and its solution within 1e-10."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _w2_mass(hip, orc, n, nl, seed=0):
    w2 = hip.FunctionSpace.cubedsphere(n, nl, hip.W2, 0)
    M2 = hip.LmaOperator(w2).generate_spd(seed, shift=1.0)
    m = orc.Mesh(n, nl, orc.W2, 0)
    return w2, M2, m, orc.gen_spd(6, m.ncell_3d, seed, shift=1.0)


@pytest.mark.parametrize("precision,tol", [(0, 1e-13), (1, 2e-6)])
def test_dg_inc_apply_parity(hip, orc, precision, tol):
    w2 = hip.FunctionSpace.cubedsphere(6, 5, hip.W2, 0)
    w3 = w2.sibling(hip.W3, 0)
    D = hip.LmaOperator(w3, w2).generate_uniform(0, 5, 0.5, 1.5, orc.D_SIGNS)
    o = orc.Helmholtz(6, 5)
    if precision:
        D = D.convert(hip.R32)
    rng = np.random.default_rng(3)
    x2, y0 = rng.standard_normal(w2.undf), rng.standard_normal(w3.undf)
    if precision:
        x2, y0 = x2.astype(np.float32).astype(np.float64), y0.astype(np.float32).astype(np.float64)
    y = hip.FieldVector(w3, precision).import_field(y0)
    D.apply_dg_inc(hip.FieldVector(w2, precision).import_field(x2), y)
    Do = o.D.astype(np.float32).astype(np.float64) if precision else o.D
    ref = orc.dg_inc_apply(o.w3, o.w2, Do, x2, y0)
    scale = orc.dg_apply(o.w3, o.w2, np.abs(Do), np.abs(x2)) + np.abs(y0)
    assert np.linalg.norm(y.export_field() - ref) <= tol * np.linalg.norm(scale)
    # accumulates: a second call adds D x again
    D.apply_dg_inc(hip.FieldVector(w2, precision).import_field(x2), y)
    ref2 = orc.dg_inc_apply(o.w3, o.w2, Do, x2, ref)
    assert np.linalg.norm(y.export_field() - ref2) <= 2 * tol * np.linalg.norm(scale)


def test_boundary_dofs_and_enforce_bc_bit_exact(hip, orc):
    w2, _, m, _ = _w2_mass(hip, orc, 5, 4)
    bv = w2.boundary_dofs()
    np.testing.assert_array_equal(bv, orc.boundary_dofs(orc.W2, 0, 6))
    f = np.random.default_rng(0).standard_normal(w2.undf)
    for precision in (0, 1):
        fh = f.astype(np.float32).astype(np.float64) if precision else f
        v = hip.FieldVector(w2, precision).import_field(fh).enforce_bc()
        np.testing.assert_array_equal(v.export_field(), orc.enforce_bc(m, bv, fh))
    # W0 and W3 flag nothing: enforce_bc leaves the field as it is
    for sp in (hip.W0, hip.W3):
        s = w2.sibling(sp, 0)
        assert s.boundary_dofs().min() == 1
        g = np.random.default_rng(1).standard_normal(s.undf)
        np.testing.assert_array_equal(hip.FieldVector(s).import_field(g).enforce_bc().export_field(), g)
    # a caller's flags: W0 bottom vertices (dofs 1-4) at the bottom, top vertices (5-8) at the top
    w0 = hip.FunctionSpace.cubedsphere(5, 4)
    m0 = orc.Mesh(5, 4)
    fb = np.ones((8, 2), np.int32)
    fb[:4, 0] = 0
    fb[4:, 1] = 0
    w0.set_boundary_dofs(fb)
    np.testing.assert_array_equal(w0.boundary_dofs(), fb)
    g = np.random.default_rng(2).standard_normal(w0.undf)
    out = hip.FieldVector(w0).import_field(g).enforce_bc().export_field()
    np.testing.assert_array_equal(out, orc.enforce_bc(m0, fb, g))
    assert np.count_nonzero(out == 0.0) == 2 * m0.nverts


@pytest.mark.parametrize("form", [0, 1])
def test_mass_matrix_operator_with_bc(hip, orc, form):
    w2, M2, m, A = _w2_mass(hip, orc, 6, 5)
    bv = orc.boundary_dofs(orc.W2, 0, 6)
    x = np.random.default_rng(4).standard_normal(w2.undf)
    hip.set_apply_form(form)
    try:
        for apply_bc in (True, False):
            L = hip.LinearOperator.mass_matrix(M2, apply_bc)
            y = hip.FieldVector(w2)
            L.apply(hip.FieldVector(w2).import_field(x), y)
            ref = orc.lma_apply(m, A, x)
            if apply_bc:
                ref = orc.enforce_bc(m, bv, ref)
            yg = y.export_field()
            assert np.linalg.norm(yg - ref) <= 1e-13 * np.linalg.norm(ref)
            assert np.all(yg[ref == 0.0] == 0.0)
    finally:
        hip.set_apply_form(1)


@pytest.mark.parametrize("precision", [0, 1])
def test_w2_mass_solve_with_bc_matches_oracle(hip, orc, precision):
    """The W2 mass-matrix solve as mass_matrix_solver_alg sets it up for W2 (apply_bc on,
    diagonal preconditioner, CG): the right-hand side has the boundary conditions applied, the
    iteration stays in the interior subspace and solves M_II x_I = b_I."""
    w2, M2, m, A = _w2_mass(hip, orc, 8, 6)
    bv = orc.boundary_dofs(orc.W2, 0, 6)
    b = orc.enforce_bc(m, bv, orc.gen_uniform(m.undf, 0, 3))
    diag = orc.mm_diagonal(m, A)
    if precision:
        M2 = M2.convert(hip.R32)
    L = hip.LinearOperator.mass_matrix(M2, True)
    prec = hip.DiagonalPreconditioner.from_operator(M2)
    x = hip.FieldVector(w2, precision)
    x.set_scalar(0.0)
    bh = hip.FieldVector(w2, precision).import_field(b)
    r_tol = 1e-10 if precision == 0 else 1e-5
    res = hip.solve("cg", L, prec, x, bh, r_tol=r_tol, a_tol=0.0, max_iter=500, check_every=4)
    assert res.status == 0
    xg = x.export_field()
    surf = np.concatenate([m.dofmap[:, 4] - 1, m.dofmap[:, 4] - 1 + m.nlayers])
    assert np.all(xg[surf] == 0.0)
    if precision == 0:
        xo, reso, _ = orc.solve_callbacks("cg", m.undf, lambda v: orc.enforce_bc(m, bv, orc.lma_apply(m, A, v)),
                                          lambda v: v / diag, b, r_tol=r_tol, a_tol=0.0, max_iter=500)
        assert res.iterations == reso.iterations
        assert np.linalg.norm(xg - xo) <= 1e-10 * np.linalg.norm(xo)
    # the interior block solved directly
    S = orc.assembled_csr(m, A).tocsr()
    interior = np.setdiff1d(np.arange(m.undf), surf)
    import scipy.sparse.linalg as spla
    xi = spla.spsolve(S[interior][:, interior].tocsc(), b[interior])
    tol = 1e-8 if precision == 0 else 1e-3
    assert np.linalg.norm(xg[interior] - xi) <= tol * np.linalg.norm(xi)
