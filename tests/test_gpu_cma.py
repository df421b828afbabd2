"""Columnwise (CMA) operators and the vertical tridiagonal solve on the GPU, against the oracle
(itself bit-identical to the reference's CMA kernels compiled from source).

Bars: every CMA kernel (asm, app, appinv, mul, scaledadd, asm_diag_hmht, asm_m2v_lumped_inv)
and tri_solve equal to the oracle bit for bit, FP32 (r_solver) and FP64, on a box mesh of W3,
Wtheta-like and W2h spaces; the vertical-line Helmholtz preconditioner built from them (M3 +
tau (Dv M2v^-1 Dv^T + diag(Dh M2h^-1 Dh^T)), the reference's columnwise construction) cuts the
GMRES residual of the assembled W2/W3 Helmholtz system on thin layers more than 20-fold per cycle
and the preconditioned solve meets the tolerance."""
import numpy as np
import pytest

from test_gpu_assembly import _box_mesh, _box_w2_w3, _chi_of_box

pytestmark = pytest.mark.gpu


def _spaces(hip, nx, ny, nl):
    """W3 and a Wtheta-like space (bottom and top face per cell) on one box mesh with one
    colouring, with their dofmaps."""
    _, npc, cic, _ = _box_mesh(nx, ny, nl)
    _, _, dm3, undf3 = _box_w2_w3(nx, ny, nl)
    nc = nx * ny
    dmt = np.array([[c * (nl + 1) + 1, c * (nl + 1) + 2] for c in range(nc)], np.int32)
    w3 = hip.FunctionSpace.from_dofmap(nl, undf3, dm3, npc, cic)
    wt = hip.FunctionSpace.from_dofmap(nl, nc * (nl + 1), dmt, npc, cic)
    return w3, dm3, wt, dmt


def _orc_space(orc, nl, dm, ni, nf):
    banded, ind = orc.cma_dofmaps(nl, dm.shape[1], ni, nf, dm[0])
    return banded, ind


def _upload_random(hip, rng, to, fr, P, lo=-1.0, hi=1.0):
    a = rng.uniform(lo, hi, (to.ncell_3d, to.ndf, fr.ndf)).astype(np.float32).astype(np.float64)
    return hip.LmaOperator(to, fr, P).upload(a), a


@pytest.mark.parametrize("P", ["R32", "R64"])
def test_cma_kernels_bitwise_equal_oracle(hip, orc, P):
    P = getattr(hip, P)
    r32 = P == hip.R32
    T = np.float32 if r32 else np.float64
    nx, ny, nl = 6, 5, 5
    nc = nx * ny
    rng = np.random.default_rng(11)
    w3, dm3, wt, dmt = _spaces(hip, nx, ny, nl)
    b3, i3 = _orc_space(orc, nl, dm3, 1, 0)
    bt, it = _orc_space(orc, nl, dmt, 0, 1)
    sp = {id(w3): (dm3, b3, i3, (1, 0)), id(wt): (dmt, bt, it, (0, 1))}
    cmas = {}
    for to, fr in ((w3, wt), (wt, w3), (wt, wt), (w3, w3)):
        print("assemble", to.ndf, fr.ndf, flush=True)
        lma, a = _upload_random(hip, rng, to, fr, P)
        A = hip.ColumnwiseOperator(to, fr, P).assemble(lma)
        (mt, bnt, indt, (nit, nft)), (mf, bnf, indf, (nif, nff)) = sp[id(to)], sp[id(fr)]
        p = orc.cma_params(nl, nit, nft, nif, nff)
        np.testing.assert_array_equal(A.params, p)
        ref = orc.cma_asm(nc, nl, a, p, bnt, bnf, r32=r32)
        np.testing.assert_array_equal(A.download(), ref)
        x = rng.uniform(-1, 1, fr.undf).astype(T)
        y0 = rng.uniform(-1, 1, to.undf).astype(T)
        xv = hip.FieldVector(fr, P).import_raw(x)
        yv = hip.FieldVector(to, P).import_raw(y0)
        A.apply(xv, yv, accumulate=True)
        np.testing.assert_array_equal(yv.export_raw(), orc.cma_apply(p, ref, x, y0, mt, indt, mf, indf, r32=r32))
        A.apply(xv, yv)
        np.testing.assert_array_equal(yv.export_raw(),
                                      orc.cma_apply(p, ref, x, np.zeros_like(y0), mt, indt, mf, indf, r32=r32))
        cmas[(id(to), id(fr))] = (A, p, ref)
    # C = A B (onto a nonzero C), S = 2.4 M + 3.7 C, and the tridiagonal solve of S + 8 I
    A, pa, ra = cmas[(id(w3), id(wt))]
    B, pb, rb = cmas[(id(wt), id(w3))]
    M, pm, rm = cmas[(id(w3), id(w3))]
    C = hip.ColumnwiseOperator(w3, w3, P, custom=(1, 1, 1, 1))
    pc = orc.cma_params(nl, 1, 0, 1, 0, custom=(1, 1, 1, 1))
    c0 = rng.uniform(-1, 1, (nc, pc[5], pc[4])).astype(np.float32).astype(np.float64)
    C.upload(c0)
    A.multiply_into(B, C)
    rc = orc.cma_mul(pa, ra, pb, rb, pc, c0.astype(T), r32=r32)
    np.testing.assert_array_equal(C.download(), rc)
    S = hip.ColumnwiseOperator(w3, w3, P, custom=(1, 1, 1, 1))
    M.scaled_add_into(C, S, 2.4, 3.7)
    alpha, beta = (np.float32(2.4), np.float32(3.7)) if r32 else (2.4, 3.7)
    rs = orc.cma_scaledadd(pm, rm, pc, rc, pc, alpha, beta, r32=r32)
    np.testing.assert_array_equal(S.download(), rs)
    rs[:, :, 1] += T(8.0)
    S.upload(rs)
    b = rng.uniform(-1, 1, w3.undf).astype(T)
    z = hip.FieldVector(w3, P)
    S.apply_inverse(hip.FieldVector(w3, P).import_raw(b), z)
    np.testing.assert_array_equal(z.export_raw(), orc.cma_appinv(pc, rs, b, dm3, i3, r32=r32))
    # lumped inverse of a Wtheta operator and Dh diag(M2h)^-1 Dh^T with a 4-dof side space
    lma, a = _upload_random(hip, rng, wt, wt, P, 0.5, 1.5)
    L = hip.ColumnwiseOperator(wt, wt, P, custom=(1, 1, 0, 0)).assemble_m2v_lumped_inv(lma)
    pl = orc.cma_params(nl, 0, 1, 0, 1, custom=(1, 1, 0, 0))
    np.testing.assert_array_equal(L.download(), orc.cma_asm_m2v_lumped_inv(nc, nl, a, pl, bt, r32=r32))
    # tri_solve on W3 fields
    t0 = rng.uniform(3, 5, w3.undf).astype(T)
    tp, tm, x = (rng.uniform(-1, 1, w3.undf).astype(T) for _ in range(3))
    f = lambda v: hip.FieldVector(w3, P).import_raw(v)
    y, work = hip.FieldVector(w3, P), hip.FieldVector(w3, P)
    hip.tri_solve(f(x), f(t0), f(tp), f(tm), y, work)
    np.testing.assert_array_equal(y.export_raw(), orc.tri_solve(nl, dm3, x, t0, tp, tm, r32=r32))


@pytest.mark.parametrize("P", ["R32", "R64"])
@pytest.mark.parametrize("nl", [100, 140])
def test_cma_tall_columns_bitwise_equal_oracle(hip, orc, P, nl):
    """Columns taller than one chunk of rows (100: the LDS-staged apply and solve take two or
    three chunks of bands) and taller than the LDS-staged kernels accept (140: the kernels with
    one thread per column and no staging), over 70 columns (two blocks of 64, the second
    partial): apply, accumulate and the tridiagonal solve equal the oracle bit for bit."""
    P = getattr(hip, P)
    r32 = P == hip.R32
    T = np.float32 if r32 else np.float64
    nx, ny = 10, 7
    nc = nx * ny
    rng = np.random.default_rng(nl)
    print("tall columns", nl, "rows", "R32" if r32 else "R64", flush=True)
    w3, dm3, _, _ = _spaces(hip, nx, ny, nl)
    _, i3 = _orc_space(orc, nl, dm3, 1, 0)
    pc = orc.cma_params(nl, 1, 0, 1, 0, custom=(1, 1, 1, 1))
    S = hip.ColumnwiseOperator(w3, w3, P, custom=(1, 1, 1, 1))
    bands = rng.uniform(-1, 1, (nc, pc[5], pc[4])).astype(np.float32).astype(np.float64)
    bands[:, :, 1] += 4.0
    S.upload(bands)
    ref = S.download()
    x = rng.uniform(-1, 1, w3.undf).astype(T)
    y0 = rng.uniform(-1, 1, w3.undf).astype(T)
    xv = hip.FieldVector(w3, P).import_raw(x)
    y = hip.FieldVector(w3, P)
    S.apply(xv, y)
    np.testing.assert_array_equal(y.export_raw(),
                                  orc.cma_apply(pc, ref, x, np.zeros_like(y0), dm3, i3, dm3, i3, r32=r32))
    y.import_raw(y0)
    S.apply(xv, y, accumulate=True)
    np.testing.assert_array_equal(y.export_raw(), orc.cma_apply(pc, ref, x, y0, dm3, i3, dm3, i3, r32=r32))
    S.apply_inverse(xv, y)
    np.testing.assert_array_equal(y.export_raw(), orc.cma_appinv(pc, ref, x, dm3, i3, r32=r32))


def test_cma_diag_hmht_bitwise_equals_oracle(hip, orc):
    """asm_diag_hmht from a W3 <- W2h LMA operator and a W2h mass (the side faces of the box's W2
    dofmap, shared between neighbouring cells)."""
    nx, ny, nl = 4, 3, 4
    nc = nx * ny
    _, npc, cic, _ = _box_mesh(nx, ny, nl)
    dm2, undf2, dm3, undf3 = _box_w2_w3(nx, ny, nl)
    w3 = hip.FunctionSpace.from_dofmap(nl, undf3, dm3, npc, cic)
    side = dm2[:, :4]  # x- and y-normal face columns come first in the W2 numbering
    w2h = hip.FunctionSpace.from_dofmap(nl, int(side.max()) + nl - 1, side, npc, cic)
    rng = np.random.default_rng(12)
    for P, r32 in ((hip.R32, True), (hip.R64, False)):
        T = np.float32 if r32 else np.float64
        d = rng.uniform(-1, 1, (nc * nl, 1, 4)).astype(np.float32).astype(np.float64)
        m = (rng.uniform(-1, 1, (nc * nl, 4, 4)) + 4 * np.eye(4)).astype(np.float32).astype(np.float64)
        Dh = hip.LmaOperator(w3, w2h, P).upload(d)
        M2h = hip.LmaOperator(w2h, w2h, P).upload(m)
        H = hip.ColumnwiseOperator(w3, w3, P).assemble_diag_hmht(Dh, M2h)
        b3, _ = orc.cma_dofmaps(nl, 1, 1, 0, dm3[0])
        np.testing.assert_array_equal(H.download(),
                                      orc.cma_asm_diag_hmht(nc, nl, d.astype(T), m.astype(T), H.params, b3, r32=r32))


def test_cma_vertical_helmholtz_preconditioner(hip, orc):
    """The columnwise Helmholtz preconditioner of the assembled W2/W3 system H = M3 + tau D
    diag(M2)^-1 D^T on a box: Hc = M3 + tau (Dv M2v^-1 Dv^T + diag(Dh M2h^-1 Dh^T)) from the
    vertical (bottom/top) and side parts of the assembled divergence and W2 mass, solved column
    by column.  On thin (atmosphere-like) layers one GMRES cycle with it leaves a residual more
    than 20 times below one without, and the preconditioned solve converges."""
    nx, ny, nl, tau = 8, 7, 12, 40.0
    nc = nx * ny
    dm0, npc, cic, undf0 = _box_mesh(nx, ny, nl)
    dm2, undf2, dm3, undf3 = _box_w2_w3(nx, ny, nl)
    w0 = hip.FunctionSpace.from_dofmap(nl, undf0, dm0, npc, cic)
    w2 = hip.FunctionSpace.from_dofmap(nl, undf2, dm2, npc, cic)
    w3 = hip.FunctionSpace.from_dofmap(nl, undf3, dm3, npc, cic)
    # the vertical faces as a Wtheta-like space (bottom, top) and the side faces as W2h
    zb = dm2[:, 4] - dm2[:, 4].min()  # horizontal-face columns, renumbered from 0
    w2v_map = np.stack([zb + 1, zb + 2], axis=1).astype(np.int32)
    w2v = hip.FunctionSpace.from_dofmap(nl, nc * (nl + 1), w2v_map, npc, cic)
    side = dm2[:, :4]
    w2h = hip.FunctionSpace.from_dofmap(nl, int(side.max()) + nl - 1, side, npc, cic)
    chi = _chi_of_box(nx, ny, nl, np.random.default_rng(13))
    chi[2] *= 0.02  # thin layers, as in an atmosphere: the vertical coupling dominates
    cv = [hip.FieldVector(w0).import_field(c) for c in chi]
    M2 = hip.LmaOperator(w2).assemble_mass(*cv)
    M3 = hip.LmaOperator(w3).assemble_mass(*cv)
    D = hip.LmaOperator(w3, w2).assemble_div(*cv)
    H = hip.LinearOperator.helmholtz(M3, D, M2.mm_diagonal(hip.FieldVector(w2)), tau)
    # split the element matrices into their vertical and side parts (set-up on the host)
    d, m2 = D.download(), M2.download()
    Dv = hip.LmaOperator(w3, w2v).upload(d[:, :, 4:6])
    DvT = hip.LmaOperator(w2v, w3).upload(d[:, :, 4:6].transpose(0, 2, 1))
    M2v = hip.LmaOperator(w2v, w2v).upload(m2[:, 4:6, 4:6])
    Dh = hip.LmaOperator(w3, w2h).upload(d[:, :, 0:4])
    M2h = hip.LmaOperator(w2h, w2h).upload(m2[:, 0:4, 0:4])
    R64 = hip.R64
    m3c = hip.ColumnwiseOperator(w3, w3, R64).assemble(M3)
    dvc = hip.ColumnwiseOperator(w3, w2v, R64).assemble(Dv)
    dvtc = hip.ColumnwiseOperator(w2v, w3, R64).assemble(DvT)
    minv = hip.ColumnwiseOperator(w2v, w2v, R64, custom=(1, 1, 0, 0)).assemble_m2v_lumped_inv(M2v)
    hh = hip.ColumnwiseOperator(w3, w3, R64).assemble_diag_hmht(Dh, M2h)
    t1 = hip.ColumnwiseOperator(w2v, w3, R64, custom=tuple(dvtc.params[:4]))
    minv.multiply_into(dvtc, t1)                      # M2v^-1 Dv^T
    t2 = hip.ColumnwiseOperator(w3, w3, R64, custom=(1, 1, 1, 1))
    dvc.multiply_into(t1, t2)                         # Dv M2v^-1 Dv^T (tridiagonal)
    vert = hip.ColumnwiseOperator(w3, w3, R64, custom=(1, 1, 1, 1))
    m3c.scaled_add_into(t2, vert, 1.0, tau)
    hc = hip.ColumnwiseOperator(w3, w3, R64, custom=(1, 1, 1, 1))
    vert.scaled_add_into(hh, hc, 1.0, tau)
    b = hip.FieldVector(w3).fill_uniform(14, 3)
    bb = b.export_field()

    def true_res(x):
        y = hip.FieldVector(w3)
        H.apply(x, y)
        return np.linalg.norm(y.export_field() - bb) / np.linalg.norm(bb)

    # one GMRES cycle of 10 iterations with each preconditioner (max_iter counts restart cycles
    # and GMRES tests convergence only at the end of a cycle, as the reference's does)
    res = {}
    precs = (("none", hip.NullPreconditioner()), ("columnwise", hip.ColumnwiseInversePreconditioner(hc)))
    for name, prec in precs:
        x = hip.FieldVector(w3)
        hip.GMRES(H, prec, 10, 1e-14, 0.0, 1).apply(x, b)
        res[name] = true_res(x)
    print("relative residual after 10 GMRES iterations", res, flush=True)
    assert res["columnwise"] < 0.05 * res["none"]
    # and the preconditioned solve converges to the tolerance
    x = hip.FieldVector(w3)
    r = hip.GMRES(H, precs[1][1], 10, 1e-10, 0.0, 20).apply(x, b)
    assert r.status == 0 and true_res(x) <= 1e-9, r
