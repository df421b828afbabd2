"""GPU parity of the mixed W2/W3 path: transpose and DG applies, the Helmholtz-like composite
operator, the vertical-column block-Jacobi preconditioner and GMRES.

Bars: FP64 operator applies within 1e-13 (norm-wise relative) of the oracle composition;
column block-Jacobi within 2e-5 (FP32 factors) / 3e-2 (BF16 factors) of exact FP64 tridiagonal
column solves; FP64 GMRES with no preconditioner matches the oracle's restart count and
solution to 1e-9; mixed-precision GMRES reaches a true FP64 relative residual <= 1e-8."""
import numpy as np
import pytest
import scipy.linalg as sla

pytestmark = pytest.mark.gpu


def _helmholtz(hip, orc, n, nl, precision=0, tau=0.5, seed=0):
    w2 = hip.FunctionSpace.cubedsphere(n, nl, hip.W2, 0)
    w3 = w2.sibling(hip.W3, 0)
    M2 = hip.LmaOperator(w2).generate_spd(seed, shift=1.0)
    D = hip.LmaOperator(w3, w2).generate_uniform(seed, 5, 0.5, 1.5, orc.D_SIGNS)
    M3 = hip.LmaOperator(w3).generate_spd(seed + 1, shift=1.0)
    if precision != 0:
        M2, D, M3 = M2.convert(precision), D.convert(precision), M3.convert(precision)
    m2d = M2.mm_diagonal(hip.FieldVector(w2, precision))
    H = hip.LinearOperator.helmholtz(M3, D, m2d, tau)
    return w2, w3, H, (M2, D, M3, m2d)


def test_generated_mixed_operators_match_oracle(hip, orc):
    w2, w3, H, (M2, D, M3, m2d) = _helmholtz(hip, orc, 4, 3)
    o = orc.Helmholtz(4, 3)
    np.testing.assert_array_equal(D.download(), o.D)
    np.testing.assert_array_equal(M2.download(), o.M2)
    np.testing.assert_array_equal(M3.download(), o.M3)
    np.testing.assert_allclose(m2d.export_field(), o.m2d, rtol=1e-15)


def test_transpose_and_dg_apply_parity(hip, orc):
    w2, w3, H, (M2, D, M3, m2d) = _helmholtz(hip, orc, 12, 7)
    o = orc.Helmholtz(12, 7)
    rng = np.random.default_rng(0)
    x3 = rng.standard_normal(w3.undf)
    x2 = rng.standard_normal(w2.undf)
    y2 = hip.FieldVector(w2)
    D.apply_transpose(hip.FieldVector(w3).import_field(x3), y2)
    ref = orc.transpose_apply(o.w2, o.w3, o.D, x3)
    assert np.linalg.norm(y2.export_field() - ref) <= 1e-13 * np.linalg.norm(ref)
    y3 = hip.FieldVector(w3)
    D.apply_dg(hip.FieldVector(w2).import_field(x2), y3)
    ref = orc.dg_apply(o.w3, o.w2, o.D, x2)
    assert np.linalg.norm(y3.export_field() - ref) <= 1e-13 * np.linalg.norm(ref)
    # W0 transpose of a non-symmetric operator = apply of its transpose
    fs = hip.FunctionSpace.cubedsphere(4, 3)
    m = orc.Mesh(4, 3)
    A = rng.standard_normal((m.ncell_3d, 8, 8))
    op = hip.LmaOperator(fs).upload(A)
    x = rng.standard_normal(m.undf)
    y = hip.FieldVector(fs)
    op.apply_transpose(hip.FieldVector(fs).import_field(x), y)
    ref = orc.lma_apply(m, A.transpose(0, 2, 1), x)
    assert np.linalg.norm(y.export_field() - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("n,nl", [(4, 3), (12, 10)])
def test_helmholtz_apply_parity(hip, orc, n, nl):
    w2, w3, H, _ = _helmholtz(hip, orc, n, nl)
    o = orc.Helmholtz(n, nl)
    x = np.random.default_rng(1).standard_normal(w3.undf)
    y = hip.FieldVector(w3)
    H.apply(hip.FieldVector(w3).import_field(x), y)
    ref = o.apply(x)
    assert np.linalg.norm(y.export_field() - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n,nl", [(4, 3), (5, 7), (12, 10)])
def test_helmholtz_two_kernel_apply_equals_the_kernel_chain(hip, orc, n, nl, precision):
    """The Helmholtz composite's two-kernel apply (transpose in gather form with the division,
    over the composite's copy of D in transpose-gather order, then divergence, M3 and the axpy
    in one kernel) equals the chain of separate applies - colour-ordered transpose, division by
    diag(M2), divergence, M3, y += tau t - bit for bit; after D's or M3's entries are
    rewritten the composite's copies follow them."""
    tau = 0.5
    w2, w3, H, (M2, D, M3, m2d) = _helmholtz(hip, orc, n, nl, precision, tau)
    x = hip.FieldVector(w3, precision).fill_uniform(3, 3)
    for rewrite in (None, D, M3):
        if rewrite is not None:
            rewrite.upload(np.roll(rewrite.download(), 1, axis=0))
        y = hip.FieldVector(w3, precision)
        H.apply(x, y)
        t2, t2d = hip.FieldVector(w2, precision), hip.FieldVector(w2, precision)
        t3, yc = hip.FieldVector(w3, precision), hip.FieldVector(w3, precision)
        D.apply_transpose(x, t2)
        t2d.divide(t2, m2d)
        D.apply_dg(t2d, t3)
        M3.apply_dg(x, yc)
        yc.axpy(tau, t3)
        np.testing.assert_array_equal(y.export_raw(), yc.export_raw())


def _ir_gmres_columns(apply, b, m, r_tol, inner_tol, max_outer=50):
    """FP64 numpy restatement of the refinement around GMRES(m) whose inner solve ends its
    cycle on the least-squares residual estimate: the estimate of column iv-1 is looked at once
    column iv is made (and only while iv < m-1), and the cycle then keeps iv+1 columns; a cycle
    of m columns ends with the inner true-residual test.  Returns (columns, outer steps)."""
    x = np.zeros_like(b)
    r = b.copy()
    nr0 = nr = np.linalg.norm(r)
    total = outer = 0
    while nr / nr0 >= r_tol and outer < max_outer:
        rhs = r / nr
        tol = max(inner_tol, 0.5 * r_tol * nr0 / nr)
        d = np.zeros_like(b)
        beta0 = None
        while True:
            res = rhs - apply(d)
            beta = np.linalg.norm(res)
            beta0 = beta if beta0 is None else beta0
            V, H, g = [res / beta], np.zeros((m + 1, m)), np.zeros(m + 1)
            g[0] = beta
            cs, sn, ratios, k = np.zeros(m), np.zeros(m), [], m
            for iv in range(m):
                w = apply(V[iv])
                for j in range(iv + 1):
                    H[j, iv] = V[j] @ w
                    w = w - H[j, iv] * V[j]
                H[iv + 1, iv] = np.linalg.norm(w)
                V.append(w / H[iv + 1, iv])
                for i in range(iv):
                    a, c = H[i, iv], H[i + 1, iv]
                    H[i, iv], H[i + 1, iv] = cs[i] * a + sn[i] * c, -sn[i] * a + cs[i] * c
                nrm = np.hypot(H[iv, iv], H[iv + 1, iv])
                cs[iv], sn[iv] = H[iv, iv] / nrm, H[iv + 1, iv] / nrm
                g[iv], g[iv + 1] = cs[iv] * g[iv] + sn[iv] * g[iv + 1], -sn[iv] * g[iv] + cs[iv] * g[iv + 1]
                H[iv, iv] = cs[iv] * H[iv, iv] + sn[iv] * H[iv + 1, iv]
                ratios.append(abs(g[iv + 1]) / beta0)
                if 1 <= iv < m - 1 and ratios[iv - 1] < tol:
                    k = iv + 1
                    break
            d = d + np.column_stack(V[:k]) @ np.linalg.solve(np.triu(H[:k, :k]), g[:k])
            total += k
            if k < m or np.linalg.norm(rhs - apply(d)) < tol * np.linalg.norm(rhs):
                break
        x = x + nr * d
        r = b - apply(x)
        nr = np.linalg.norm(r)
        outer += 1
    return total, outer


@pytest.mark.parametrize("gcrk", [30, 6])
def test_mixed_gmres_inner_exit_matches_restatement(hip, orc, gcrk):
    """The inner solves' exits on the residual estimate (no preconditioner, so the estimate is
    the residual's): as many refinement steps as the FP64 restatement, and the same number of
    Arnoldi columns within one per step (FP32 inner arithmetic against FP64).  gcrk 6 needs
    whole restarted cycles before the exit."""
    w2, w3, H64, (M2, D, M3, m2d) = _helmholtz(hip, orc, 12, 10)
    M2s, Ds, M3s = M2.convert(hip.R32), D.convert(hip.R32), M3.convert(hip.R32)
    H32 = hip.LinearOperator.helmholtz(M3s, Ds, M2s.mm_diagonal(hip.FieldVector(w2, hip.R32)), 0.5)
    o = orc.Helmholtz(12, 10)
    b = np.random.default_rng(5).uniform(-1, 1, w3.undf)
    x = hip.FieldVector(w3)
    res = hip.MixedPrecisionGMRES(H64, H32, hip.NullPreconditioner(), gcrk, 1e-8, 0.0, 50,
                                  inner_r_tol=1e-4).apply(x, hip.FieldVector(w3).import_field(b))
    cols, outer = _ir_gmres_columns(o.apply, b, gcrk, 1e-8, 1e-4)
    print(f"gcrk {gcrk}: device {res.iterations} columns / {res.outer_iterations} steps, "
          f"restatement {cols} / {outer}", flush=True)
    assert res.status == 0 and res.outer_iterations == outer
    assert abs(res.iterations - cols) <= outer


def _column_solves(Hd, columns, r):
    z = np.zeros_like(r)
    for base, h in columns:
        idx = np.arange(base, base + h)
        T = Hd[idx][:, idx]
        T = T.toarray() if hasattr(T, "toarray") else T
        ab = np.zeros((3, h))
        ab[0, 1:] = np.diag(T, 1)
        ab[1] = np.diag(T)
        ab[2, :-1] = np.diag(T, -1)
        z[idx] = sla.solve_banded((1, 1), ab, r[idx])
    return z


@pytest.mark.parametrize("storage,tol", [(1, 2e-5), (2, 3e-2), (3, 5e-3)])
def test_column_jacobi_helmholtz(hip, orc, storage, tol):
    w2, w3, H, _ = _helmholtz(hip, orc, 4, 5)
    o = orc.Helmholtz(4, 5)
    Hd = o.dense()
    cols = [(o.w3.dofmap[c, 0] - 1, 5) for c in range(o.w3.ncells)]
    P = hip.ColumnJacobiPreconditioner(H, storage)
    r = np.random.default_rng(2).standard_normal(w3.undf)
    z = hip.FieldVector(w3)
    P.apply(hip.FieldVector(w3).import_field(r), z)
    ref = _column_solves(Hd, cols, r)
    assert np.linalg.norm(z.export_field() - ref) <= tol * np.linalg.norm(ref)


@pytest.mark.parametrize("nl", [6, 120])
def test_column_jacobi_w0_lma(hip, orc, nl):
    """W0 vertex columns of nl+1 dofs; at 120 levels fewer than 32 columns fit one block's LDS
    tile (FP32 factors: 25 per block)."""
    fs = hip.FunctionSpace.cubedsphere(4, nl)
    m = orc.Mesh(4, nl)
    op = hip.LmaOperator(fs).generate_spd(3)
    A = orc.gen_spd(8, m.ncell_3d, 3)
    S = orc.assembled_csr(m, A).tocsr()
    bases = sorted(set(m.dofmap[:, :4].ravel() - 1))
    P = hip.ColumnJacobiPreconditioner(hip.LinearOperator.from_lma(op), hip.R32)
    r = np.random.default_rng(3).standard_normal(m.undf)
    z = hip.FieldVector(fs)
    P.apply(hip.FieldVector(fs).import_field(r), z)
    ref = _column_solves(S, [(b, nl + 1) for b in bases], r)
    assert np.linalg.norm(z.export_field() - ref) <= 2e-5 * np.linalg.norm(ref)


def test_gmres_fp64_matches_oracle(hip, orc):
    w2, w3, H, _ = _helmholtz(hip, orc, 4, 5)
    o = orc.Helmholtz(4, 5)
    b = np.random.default_rng(4).uniform(-1, 1, w3.undf)
    x = hip.FieldVector(w3)
    solver = hip.GMRES(H, hip.NullPreconditioner(), 10, 1e-10, 0.0, 40)
    res = solver.apply(x, hip.FieldVector(w3).import_field(b))
    xo, reso, hist = orc.gmres_callbacks(w3.undf, o.apply, lambda v: v, b, gcrk=10, r_tol=1e-10, a_tol=0.0,
                                         max_iter=40)
    assert res.status == 0 and reso.converged == 1
    assert res.outer_iterations == reso.iterations
    assert np.linalg.norm(x.export_field() - xo) <= 1e-9 * np.linalg.norm(xo)


def test_mixed_gmres_reaches_fp64_accuracy(hip, orc):
    w2, w3, H64, (M2, D, M3, m2d) = _helmholtz(hip, orc, 12, 10)
    M2s, Ds, M3s = M2.convert(hip.R32), D.convert(hip.R32), M3.convert(hip.R32)
    m2ds = M2s.mm_diagonal(hip.FieldVector(w2, hip.R32))
    H32 = hip.LinearOperator.helmholtz(M3s, Ds, m2ds, 0.5)
    o = orc.Helmholtz(12, 10)
    P = hip.ColumnJacobiPreconditioner(H32, hip.BF16)
    b = np.random.default_rng(5).uniform(-1, 1, w3.undf)
    x = hip.FieldVector(w3)
    solver = hip.MixedPrecisionGMRES(H64, H32, P, 30, 1e-8, 0.0, 50, inner_r_tol=1e-4)
    res = solver.apply(x, hip.FieldVector(w3).import_field(b))
    xg = x.export_field()
    rel = np.linalg.norm(b - o.apply(xg)) / np.linalg.norm(b)
    assert res.status == 0 and rel <= 1.01e-8, (res, rel)
    # the inner solves end their cycles on the Givens residual estimate: fewer Arnoldi columns
    # than whole cycles of 30
    print(f"inner columns {res.iterations}, outer steps {res.outer_iterations}", flush=True)
    assert 0 < res.iterations < 30 * res.outer_iterations
    xo, reso, _ = orc.gmres_callbacks(w3.undf, o.apply, lambda v: v, b, gcrk=30, r_tol=1e-12, max_iter=200)
    assert np.linalg.norm(xg - xo) <= 1e-6 * np.linalg.norm(xo)
