"""GPU checks at the full size of the headline workload (C96 cubed sphere, 70 levels, W0,
3,926,158 dofs, 3,870,720 cells), where the oracle would take minutes: size-independent
properties instead of oracle values.

* the gather apply over the row-gathered copy equals the cell-layout gather bit for bit
  (FP32 and FP64), and the multi-RHS colour / matrix-core forms equal their one-RHS colour
  apply column by column;
* the operator is symmetric (x.Ay = y.Ax to FP64 rounding), as B B^T + 0.1 I element matrices
  must make it;
* the mixed-precision solve reaches an FP64 true relative residual <= 1e-8 and agrees with the
  FP64 solve to the conditioning of the system.
Small-size parity of every form against the oracle is in test_gpu_parity.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c96(hip):
    fs = hip.FunctionSpace.cubedsphere(96, 70)
    assert (fs.undf, fs.ncell_3d) == (3926158, 3870720)
    op64 = hip.LmaOperator(fs).generate_spd(0, scaled=True)
    op32 = op64.convert(hip.R32)
    return fs, op64, op32


def test_fullsize_apply_forms_agree_bitwise(hip, c96):
    fs, op64, op32 = c96
    for op, P in ((op32, hip.R32), (op64, hip.R64)):
        x = hip.FieldVector(fs, P).fill_uniform(7)
        y0, y1 = hip.FieldVector(fs, P), hip.FieldVector(fs, P)
        op.apply_gather(x, y0)
        ref = y0.export_raw()
        op.prepare_rows()
        y1.set_scalar(-3.0)
        op.apply_gather(x, y1)
        np.testing.assert_array_equal(y1.export_raw(), ref)
    # 4 right-hand sides: colour-ordered FMA and matrix-core forms against the one-RHS colour apply
    X = np.random.default_rng(3).uniform(-1, 1, (fs.undf, 4)).astype(np.float32)
    xm = hip.FieldVector(fs, hip.R32, 4).import_raw(X.ravel())
    yc, ym = hip.FieldVector(fs, hip.R32, 4), hip.FieldVector(fs, hip.R32, 4)
    op32.apply_mrhs(xm, yc, 2)
    op32.apply_mrhs(xm, ym, 3)
    Yc = yc.export_raw().reshape(fs.undf, 4)
    np.testing.assert_array_equal(ym.export_raw().reshape(fs.undf, 4), Yc)
    for r in (0, 3):
        x1, y1 = hip.FieldVector(fs, hip.R32).import_raw(X[:, r]), hip.FieldVector(fs, hip.R32)
        op32.apply(x1, y1)
        np.testing.assert_array_equal(Yc[:, r], y1.export_raw())
    # the symmetric-packed apply of a copy of the FP32 operator equals its colour-ordered apply
    ops = op64.convert(hip.R32).prepare_symmetric()
    x = hip.FieldVector(fs, hip.R32).fill_uniform(7)
    yc, ys = hip.FieldVector(fs, hip.R32), hip.FieldVector(fs, hip.R32)
    ops.apply(x, yc)
    ops.apply_gather(x, ys)
    np.testing.assert_array_equal(ys.export_raw(), yc.export_raw())


def test_fullsize_operator_is_symmetric(hip, c96):
    fs, op64, _ = c96
    x = hip.FieldVector(fs).fill_uniform(11)
    y = hip.FieldVector(fs).fill_uniform(12)
    ax, ay = hip.FieldVector(fs), hip.FieldVector(fs)
    op64.apply(x, ax)
    op64.apply(y, ay)
    a, b = y.dot(ax), x.dot(ay)
    assert abs(a - b) <= 1e-12 * max(abs(a), abs(b)), (a, b)


def test_fullsize_mixed_solve_reaches_fp64_tolerance(hip, c96):
    from lfric_core_amd.solvers import residual_norm
    fs, op64, op32 = c96
    b = hip.FieldVector(fs).fill_uniform(0, 3)
    x_mixed, x_fp64 = hip.FieldVector(fs), hip.FieldVector(fs)
    prec32 = hip.DiagonalPreconditioner.from_operator(op32)
    prec64 = hip.DiagonalPreconditioner.from_operator(op64)
    res = hip.MixedPrecisionCG(op64, op32, prec32, 1e-8, 0.0, 20000, inner_r_tol=1e-4, check_every=16).apply(x_mixed, b)
    assert res.status == 0 and res.true_rel_res <= 1e-8
    assert residual_norm(op64, x_mixed, b) / res.r_nrm_0 <= 1e-8
    res64 = hip.ConjugateGradient(op64, prec64, 1e-8, 0.0, 20000, check_every=16).apply(x_fp64, b)
    assert res64.status == 0
    xm, x6 = x_mixed.export_field(), x_fp64.export_field()
    # both residuals <= 1e-8 ||b||; the solutions differ by at most cond(A) x 2e-8 relative
    assert np.linalg.norm(xm - x6) / np.linalg.norm(x6) <= 1e-5


def test_fullsize_helmholtz_gmres_bf16_block_jacobi(hip):
    """The W2/W3 Helmholtz-like system at its full size (C48 cubed sphere, 50 levels: W2
    2,087,424 dofs, W3 691,200): GMRES(30) with FP32 applies and the BF16 vertical-column
    block-Jacobi preconditioner inside FP64 refinement reaches an FP64 true relative residual
    <= 1e-8, and agrees with the plain FP64 GMRES solution to the system's conditioning."""
    w2 = hip.FunctionSpace.cubedsphere(48, 50, hip.W2, 0)
    w3 = w2.sibling(hip.W3, 0)
    assert (w2.undf, w3.undf) == (2087424, 691200)
    from oracle.oracle import D_SIGNS
    M2 = hip.LmaOperator(w2).generate_spd(0, shift=1.0)
    D = hip.LmaOperator(w3, w2).generate_uniform(0, 5, 0.5, 1.5, D_SIGNS)
    M3 = hip.LmaOperator(w3).generate_spd(1, shift=1.0)
    H64 = hip.LinearOperator.helmholtz(M3, D, M2.mm_diagonal(hip.FieldVector(w2)), 0.5)
    M2s, Ds, M3s = M2.convert(hip.R32), D.convert(hip.R32), M3.convert(hip.R32)
    H32 = hip.LinearOperator.helmholtz(M3s, Ds, M2s.mm_diagonal(hip.FieldVector(w2, hip.R32)), 0.5)
    P = hip.ColumnJacobiPreconditioner(H32, hip.BF16)
    b = hip.FieldVector(w3).fill_uniform(9, 3)
    x = hip.FieldVector(w3)
    res = hip.MixedPrecisionGMRES(H64, H32, P, 30, 1e-8, 0.0, 50, inner_r_tol=1e-4).apply(x, b)
    assert res.status == 0 and res.true_rel_res <= 1e-8, res
    assert 0 < res.iterations < 30 * res.outer_iterations, res
    x64 = hip.FieldVector(w3)
    res64 = hip.GMRES(H64, hip.NullPreconditioner(), 30, 1e-10, 0.0, 200).apply(x64, b)
    assert res64.status == 0
    xm, xf = x.export_field(), x64.export_field()
    assert np.linalg.norm(xm - xf) <= 1e-6 * np.linalg.norm(xf)


def test_fullsize_columnwise_solve_inverts_apply(hip):
    """A tridiagonal columnwise operator on every W3 column of C96x70 (55,296 columns of 70
    rows): the Thomas solve undoes the apply (FP64 to 1e-13, FP32 to 1e-5 relative), and the
    FP32 and FP64 solves agree to FP32 rounding."""
    w3 = hip.FunctionSpace.cubedsphere(96, 70, hip.W3, 0)
    rng = np.random.default_rng(21)
    bands = rng.uniform(-1, 1, (w3.ncells, w3.nlayers, 3))
    bands[:, :, 1] += 4.0
    xs = {}
    for P, tol in ((hip.R64, 1e-13), (hip.R32, 1e-5)):
        A = hip.ColumnwiseOperator(w3, w3, P, custom=(1, 1, 1, 1)).upload(bands)
        x = hip.FieldVector(w3, P).fill_uniform(5, 3)
        y, z = hip.FieldVector(w3, P), hip.FieldVector(w3, P)
        A.apply(x, y)
        A.apply_inverse(y, z)
        xv, zv = x.export_field(), z.export_field()
        assert np.linalg.norm(zv - xv) <= tol * np.linalg.norm(xv)
        xs[P] = (y.export_field(), A)
    # the FP32 solve of the FP64 right-hand side rounded to FP32, against the FP64 solve
    y64 = xs[hip.R64][0]
    z64, z32 = hip.FieldVector(w3), hip.FieldVector(w3, hip.R32)
    xs[hip.R64][1].apply_inverse(hip.FieldVector(w3).import_field(y64), z64)
    xs[hip.R32][1].apply_inverse(hip.FieldVector(w3, hip.R32).import_field(y64), z32)
    a, b = z64.export_field(), z32.export_field()
    assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(a)
