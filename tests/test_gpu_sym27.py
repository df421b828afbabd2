"""GPU parity of the symmetric-packed order-1 W0 apply (two phases: per-unit row sums from the
upper triangles, then a colour-ordered gather).

Each row's fma chain is formed in column order from zero and the rows are added in the
colour-ordered scatter's order, so the packed apply must equal the colour-ordered apply bit for
bit: one right-hand side (lfric_lma_apply) and 2/4/8 (LFRIC_MRHS_COLOUR), FP64 and FP32.  The
oracle's apply agrees within FP64 rounding (1e-13 of |A||x| norm-wise); CG over the packed copy
converges to the oracle's solution within 1e-8 in the oracle's iteration count (+-1: the fused
dot product adds in another order).  Non-symmetric operators are refused, and rewriting an
operator with non-symmetric entries drops its copy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n,nl", [(5, 4), (6, 7), (3, 2), (12, 9)])
def test_order1_symmetric_apply_equals_colour(hip, orc, n, nl, prec):
    fs = hip.FunctionSpace.cubedsphere(n, nl, hip.W0, 1)
    m = orc.Mesh(n, nl, orc.W0, 1)
    op64 = hip.LmaOperator(fs).generate_spd(11, scaled=True)
    op = op64 if prec == 0 else op64.convert(hip.R32)
    P = hip.R64 if prec == 0 else hip.R32
    x = hip.FieldVector(fs, P).fill_uniform(4, 3)
    yc = hip.FieldVector(fs, P)
    op.apply(x, yc)
    xm, ycm = {}, {}
    for nr in (2, 4, 8):
        xm[nr] = hip.FieldVector(fs, P, nr).fill_uniform(5 + nr, 3)
        ycm[nr] = op.apply_mrhs(xm[nr], hip.FieldVector(fs, P, nr), variant=2)
    op.prepare_symmetric()
    ys = hip.FieldVector(fs, P)
    ys.set_scalar(9.0)
    op.apply_gather(x, ys)
    np.testing.assert_array_equal(ys.export_raw(), yc.export_raw())
    for nr in (2, 4, 8):
        for variant in (1, 0):
            yg = hip.FieldVector(fs, P, nr)
            yg.set_scalar(7.0)
            op.apply_mrhs(xm[nr], yg, variant=variant)
            np.testing.assert_array_equal(yg.export_raw(), ycm[nr].export_raw())
    if prec == 0:
        A = orc.gen_spd(27, m.ncell_3d, 11, scaled=True)
        xh = x.export_field()
        ref = orc.lma_apply(m, A, xh)
        scale = orc.lma_apply(m, np.abs(A), np.abs(xh))
        assert np.linalg.norm(ys.export_field() - ref) <= 1e-13 * np.linalg.norm(scale)
    # non-symmetric operators are refused; rewriting the operator with non-symmetric entries
    # drops the copy, and the gather apply is the plain one again
    nonsym = hip.LmaOperator(fs, precision=P).generate_uniform(1, 5, -1.0, 1.0)
    with pytest.raises(hip.LfricError):
        nonsym.prepare_symmetric()
    with pytest.raises(hip.LfricError):
        op64.prepare_symmetric(4)
    op.upload(nonsym.download())
    y1, y2 = hip.FieldVector(fs, P), hip.FieldVector(fs, P)
    op.apply_gather(x, y1)
    nonsym.apply_gather(x, y2)
    np.testing.assert_array_equal(y1.export_raw(), y2.export_raw())


def test_order1_symmetric_cg_matches_oracle(hip, orc):
    n, nl = 6, 5
    fs = hip.FunctionSpace.cubedsphere(n, nl, hip.W0, 1)
    m = orc.Mesh(n, nl, orc.W0, 1)
    op = hip.LmaOperator(fs).generate_spd(3).prepare_symmetric()
    b = hip.FieldVector(fs).fill_uniform(0, 3)
    x = hip.FieldVector(fs)
    x.set_scalar(0.0)
    res = hip.ConjugateGradient(op, hip.DiagonalPreconditioner.from_operator(op), 1e-8, 0.0, 2000,
                                check_every=4).apply(x, b)
    A = orc.gen_spd(27, m.ncell_3d, 3)
    xo, ro, _ = orc.cg(m, A, b.export_field(), diag=orc.mm_diagonal(m, A), r_tol=1e-8, a_tol=0.0, max_iter=2000)
    assert res.status == 0 and abs(res.iterations - ro.iterations) <= 1, (res.iterations, ro.iterations)
    assert np.linalg.norm(x.export_field() - xo) <= 1e-8 * np.linalg.norm(xo)
