"""GPU parity of the rest of the solver family (BiCGStab, FGMRES, GCR, Jacobi, Chebyshev,
precondition-only) and of the mass-matrix solver factory, against the oracle's restatements on
the same seeded operator (C4 x 3 layers, W0, A_e = B B^T + 0.1 I, Jacobi preconditioner).

Bars: FP64 iteration counts equal to the oracle's; FP64 solutions within 1e-9 (norm-wise
relative) of the oracle's for the converging methods, 1e-12 for the fixed-step Jacobi and
Chebyshev runs (no stopping decision can differ); precondition-only bit-exact (one division);
FP32 BiCGStab reaches 1e-5."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(hip, orc, n=4, nl=3, seed=0):
    fs = hip.FunctionSpace.cubedsphere(n, nl)
    m = orc.Mesh(n, nl)
    op = hip.LmaOperator(fs).generate_spd(seed)
    A = orc.gen_spd(8, m.ncell_3d, seed)
    d = orc.mm_diagonal(m, A)
    b = orc.gen_uniform(m.undf, seed, 3)
    L = hip.LinearOperator.from_lma(op)
    P = hip.DiagonalPreconditioner.from_operator(op)
    return fs, m, op, A, d, b, L, P


def _oracle(orc, method, m, A, d, b, **kw):
    return orc.solve_callbacks(method, m.undf, lambda v: orc.lma_apply(m, A, v), lambda v: v / d, b, **kw)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("method", ["bicgstab", "fgmres", "gcr"])
def test_krylov_methods_match_oracle(hip, orc, method):
    fs, m, op, A, d, b, L, P = _problem(hip, orc)
    kw = dict(r_tol=1e-10, a_tol=1e-16, max_iter=300)
    if method != "bicgstab":
        kw.update(gcrk=8, max_iter=60)
    xo, ro, _ = _oracle(orc, method, m, A, d, b, **kw)
    assert ro.converged == 1
    bv = hip.FieldVector(fs).import_field(b)
    x = hip.FieldVector(fs)
    x.set_scalar(0.0)
    cls = {"bicgstab": hip.BiCGStab, "fgmres": hip.FGMRES, "gcr": hip.GCR}[method]
    args = (L, P, kw["r_tol"], kw["a_tol"], kw["max_iter"]) if method == "bicgstab" else \
        (L, P, kw["gcrk"], kw["r_tol"], kw["a_tol"], kw["max_iter"])
    res = cls(*args, check_every=4).apply(x, bv)
    assert res.status == 0
    if method == "bicgstab":
        assert res.iterations == ro.iterations
    else:
        assert res.outer_iterations == ro.iterations
    assert _rel(x.export_field(), xo) <= 1e-9
    # the factory gives the same solve
    x2 = hip.FieldVector(fs)
    x2.set_scalar(0.0)
    r2 = hip.solve(method, L, P, x2, bv, check_every=4, **kw)
    assert r2.status == 0
    np.testing.assert_array_equal(x2.export_field(), x.export_field())


def test_jacobi_fixed_steps_and_convergence(hip, orc):
    fs, m, op, A, d, b, L, P = _problem(hip, orc)
    bv = hip.FieldVector(fs).import_field(b)
    x0 = orc.gen_uniform(m.undf, 5, 3)
    xo, ro, _ = _oracle(orc, "jacobi", m, A, d, b, x0=x0, max_iter=15, monitor=False, rho_relax=0.4)
    x = hip.FieldVector(fs).import_field(x0)
    res = hip.Jacobi(L, P, 1e-8, 1e-16, 15, monitor_convergence=False, rho_relax=0.4).apply(x, bv)
    assert res.status == 0 and res.iterations == 15
    assert _rel(x.export_field(), xo) <= 1e-12
    # monitored: same stopping iteration as the oracle
    xo, ro, _ = _oracle(orc, "jacobi", m, A, d, b, r_tol=1e-3, max_iter=4000, rho_relax=0.4)
    x = hip.FieldVector(fs)
    x.set_scalar(0.0)
    res = hip.Jacobi(L, P, 1e-3, 1e-16, 4000, rho_relax=0.4, check_every=16).apply(x, bv)
    assert ro.converged == 1 and res.status == 0 and res.iterations == ro.iterations
    assert _rel(x.export_field(), xo) <= 1e-10


def test_chebyshev_matches_oracle(hip, orc):
    fs, m, op, A, d, b, L, P = _problem(hip, orc)
    S = orc.assembled_csr(m, A).toarray()
    Dh = 1 / np.sqrt(d)
    ev = np.linalg.eigvalsh(Dh[:, None] * S * Dh[None, :])
    lmin, lmax = float(ev.min()), float(ev.max())
    # 12 steps: the residual is still far above round-off, so it can be compared
    xo, ro, _ = _oracle(orc, "chebyshev", m, A, d, b, max_iter=12, lmin=lmin, lmax=lmax)
    bv = hip.FieldVector(fs).import_field(b)
    x = hip.FieldVector(fs)
    res = hip.Chebyshev(L, P, 1e-8, 1e-16, 12, lmin=lmin, lmax=lmax).apply(x, bv)
    assert res.status == 0 and res.iterations == 12
    assert _rel(x.export_field(), xo) <= 1e-12
    assert ro.r_nrm > 1e-8 * ro.r_nrm_0
    np.testing.assert_allclose(res.r_nrm, ro.r_nrm, rtol=1e-8)
    np.testing.assert_allclose(res.r_nrm_0, ro.r_nrm_0, rtol=1e-14)
    # run to round-off, the solution still agrees
    xo, ro, _ = _oracle(orc, "chebyshev", m, A, d, b, max_iter=40, lmin=lmin, lmax=lmax)
    res = hip.Chebyshev(L, P, 1e-8, 1e-16, 40, lmin=lmin, lmax=lmax).apply(x, bv)
    assert _rel(x.export_field(), xo) <= 1e-12 and res.r_nrm <= 1e-12 * res.r_nrm_0


def test_precondition_only_is_exact(hip, orc):
    fs, m, op, A, d, b, L, P = _problem(hip, orc)
    bv = hip.FieldVector(fs).import_field(b)
    x = hip.FieldVector(fs)
    res = hip.PreconditionOnly(L, P).apply(x, bv)
    # one correctly rounded division by the device's diagonal (which equals the oracle's
    # mm_diagonal to 1e-14: its sums run in another order)
    dg = op.mm_diagonal(hip.FieldVector(fs)).export_field()
    np.testing.assert_allclose(dg, d, rtol=1e-14)
    np.testing.assert_array_equal(x.export_field(), b / dg)
    xo, ro, _ = _oracle(orc, "prec_only", m, A, d, b)
    np.testing.assert_allclose(res.r_nrm, ro.r_nrm, rtol=1e-12)


def test_bicgstab_fp32(hip, orc):
    fs, m, op, A, d, b, L, P = _problem(hip, orc)
    op32 = op.convert(hip.R32)
    L32 = hip.LinearOperator.from_lma(op32)
    P32 = hip.DiagonalPreconditioner.from_operator(op32)
    bv = hip.FieldVector(fs, hip.R32).import_field(b)
    x = hip.FieldVector(fs, hip.R32)
    x.set_scalar(0.0)
    res = hip.BiCGStab(L32, P32, 1e-5, 1e-16, 300).apply(x, bv)
    assert res.status == 0
    S = orc.assembled_csr(m, A)
    xs = x.export_field()
    assert np.linalg.norm(b - S @ xs) / np.linalg.norm(b) <= 2e-5


def test_factory_cg_and_unknown_method(hip, orc):
    fs, m, op, A, d, b, L, P = _problem(hip, orc)
    bv = hip.FieldVector(fs).import_field(b)
    x = hip.FieldVector(fs)
    x.set_scalar(0.0)
    res = hip.solve("cg", L, P, x, bv, r_tol=1e-8, max_iter=500)
    xo, ro, _ = _oracle(orc, "cg", m, A, d, b, r_tol=1e-8, max_iter=500)
    assert res.iterations == ro.iterations and _rel(x.export_field(), xo) <= 1e-10
    import ctypes
    from lfric_core_amd._lib import SolverOpts, SolverRes, lib
    o = SolverOpts()
    lib().lfric_solver_opts_default(ctypes.byref(o))
    rc = lib().lfric_solve(99, L._h, P._h, x._h, bv._h, ctypes.byref(o), ctypes.byref(SolverRes()))
    assert rc == -1 and b"unknown solver method" in lib().lfric_last_error()
