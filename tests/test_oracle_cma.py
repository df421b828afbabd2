"""The oracle's columnwise (CMA) operators and vertical tridiagonal solve, pinned by:

* the reference unit tests' known answers (columnwise_op_{asm,app,appinv,asm_diag_hmht,
  asm_m2v_lumped_inv}_kernel_mod_test.pf, sci_tri_solve_kernel_mod_test.pf);
* dense numpy restatements per column (assembly from the element matrices through the banded
  dofmaps, products, sums, solves);
* bit-level equality with the reference's own CMA kernels compiled from source
  (oracle/_ref, r_solver = real32)."""
import ctypes
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


class _Space:
    """A horizontally discontinuous lowest-order space on ncells columns of nl layers: W3
    (one interior dof) or Wtheta (bottom and top face dofs, the top of layer k the bottom of
    k + 1)."""

    def __init__(self, orc, kind, ncells, nl):
        self.ncells, self.nl = ncells, nl
        if kind == "w3":
            self.ni, self.nf, self.ndf = 1, 0, 1
            self.map = np.array([[c * nl + 1] for c in range(ncells)], np.int32)
            self.undf = ncells * nl
        else:
            self.ni, self.nf, self.ndf = 0, 1, 2
            self.map = np.array([[c * (nl + 1) + 1, c * (nl + 1) + 2] for c in range(ncells)], np.int32)
            self.undf = ncells * (nl + 1)
        self.banded, self.ind = orc.cma_dofmaps(nl, self.ndf, self.ni, self.nf, self.map[0])
        self.nrows = nl * self.ni + (nl + 1) * self.nf


def _params(orc, to, fr):
    return orc.cma_params(to.nl, to.ni, to.nf, fr.ni, fr.nf)


def _dense(p, A):
    """Dense column matrices (ncells, nrow, ncol) of banded storage: band b of row i holds
    column j = b + j_minus(i) - 1."""
    nc, nrow, bw = A.shape
    D = np.zeros((nc, p[5], p[6]))
    for i in range(1, nrow + 1):
        jm = math.ceil((p[0] * i - p[3]) / p[1])
        for b in range(1, bw + 1):
            j = b + jm - 1
            if 1 <= j <= p[6]:
                D[:, i - 1, j - 1] = A[:, i - 1, b - 1]
    return D


def _dense_from_lma(to, fr, lma):
    """Column matrices assembled straight from the element matrices through the banded dofmaps
    (independent of the band layout)."""
    nc, nl = to.ncells, to.nl
    D = np.zeros((nc, to.nrows, fr.nrows))
    bt = to.banded.reshape((nl, to.ndf))
    bf = fr.banded.reshape((nl, fr.ndf))
    for c in range(nc):
        for k in range(nl):
            for a in range(to.ndf):
                for b in range(fr.ndf):
                    D[c, bt[k, a] - 1, bf[k, b] - 1] += lma[c * nl + k, a, b]
    return D


def _column_values(sp, v):
    """(ncells, nrows) view of a field through the indirection map."""
    return np.array([[v[sp.map[c, 0] - 1 + sp.ind[i] - 1] for i in range(sp.nrows)] for c in range(sp.ncells)])


def _wtheta_mass_bands(orc, nl, ncells=1, r32=True):
    th = _Space(orc, "wtheta", ncells, nl)
    p = _params(orc, th, th)
    T = np.float32 if r32 else np.float64
    lma = np.tile(np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]], T), (ncells * nl, 1, 1))
    return th, p, lma, orc.cma_asm(ncells, nl, lma, p, th.banded, th.banded, r32=r32)


def test_cma_known_answers(orc):
    th, p, lma, A = _wtheta_mass_bands(orc, 3)
    assert list(p[:5]) == [1, 1, 1, 1, 3] and th.ndf == 2  # the tests' own size checks
    one3, one6 = np.float32(1) / np.float32(3), np.float32(1) / np.float32(6)
    nrow = p[5]
    for i in range(nrow):
        lo = one6 if i > 0 else None
        mid = one3 if i in (0, nrow - 1) else np.float32(2) / np.float32(3)
        hi = one6 if i < nrow - 1 else None
        for b, want in ((0, lo), (1, mid), (2, hi)):
            if want is not None:
                assert abs(A[0, i, b] - want) <= 1e-6
    # app and appinv against the dense tridiagonal mass matrix
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 1, th.undf).astype(np.float32)
    dense = _dense(p, A)[0]
    y = orc.cma_apply(p, A, x, np.zeros(th.undf, np.float32), th.map, th.ind, th.map, th.ind, r32=True)
    np.testing.assert_allclose(y, dense @ x, rtol=0, atol=1e-6)
    z = orc.cma_appinv(p, A, x, th.map, th.ind, r32=True)
    np.testing.assert_allclose(dense @ z, x, rtol=0, atol=1e-6)
    # lumped inverse of the same mass: 3 at the end rows, 3/2 inside
    L = orc.cma_asm_m2v_lumped_inv(1, 3, lma, p, th.banded, r32=True)
    np.testing.assert_allclose(L[0, :, 0], [3.0, 1.5, 1.5, 3.0], rtol=0, atol=1e-6)
    # diag(H M H^T) with H = (rho), M = diag(sigma): sum rho^2 / sigma on every W3 row
    w3 = _Space(orc, "w3", 1, 3)
    p33 = _params(orc, w3, w3)
    rho = np.array([2.0, 3.0, 4.0, 5.0], np.float32)
    sigma = np.array([2.0, 7.0, 9.1, 4.8], np.float32)
    Dh = np.tile(rho[None, None, :], (3, 1, 1))
    M2h = np.tile(np.diag(sigma)[None], (3, 1, 1))
    H = orc.cma_asm_diag_hmht(1, 3, Dh, M2h, p33, w3.banded, r32=True)
    answer = np.float32(0)
    for r, s in zip(rho, sigma):
        answer = answer + np.float32(r * r / s)
    np.testing.assert_allclose(H[0, :, 0], answer, rtol=0, atol=1e-6)


def test_cma_dense_restatements(orc):
    rng = np.random.default_rng(2)
    nc, nl = 5, 6
    w3, th = _Space(orc, "w3", nc, nl), _Space(orc, "wtheta", nc, nl)
    # assembly: W3 <- Wtheta (a divergence-like operator), Wtheta <- W3, Wtheta <- Wtheta, W3 <- W3
    ops = {}
    for to, fr in ((w3, th), (th, w3), (th, th), (w3, w3)):
        lma = rng.uniform(-1, 1, (nc * nl, to.ndf, fr.ndf))
        p = _params(orc, to, fr)
        A = orc.cma_asm(nc, nl, lma, p, to.banded, fr.banded)
        np.testing.assert_allclose(_dense(p, A), _dense_from_lma(to, fr, lma), rtol=1e-15, atol=1e-15)
        x = rng.uniform(-1, 1, fr.undf)
        y = orc.cma_apply(p, A, x, np.zeros(to.undf), to.map, to.ind, fr.map, fr.ind)
        want = np.einsum("cij,cj->ci", _dense(p, A), _column_values(fr, x))
        np.testing.assert_allclose(_column_values(to, y), want, rtol=1e-13, atol=1e-14)
        ops[(id(to), id(fr))] = (p, A)
    # product W3 <- Wtheta <- W3 and scaled sum with the W3 mass
    pa, A = ops[(id(w3), id(th))]
    pb, B = ops[(id(th), id(w3))]
    pm, M = ops[(id(w3), id(w3))]
    pc = orc.cma_params(nl, 1, 0, 1, 0, custom=(1, 1, 1, 1))  # tridiagonal W3 x W3
    C = orc.cma_mul(pa, A, pb, B, pc, np.zeros((nc, pc[5], pc[4])))
    np.testing.assert_allclose(_dense(pc, C), _dense(pa, A) @ _dense(pb, B), rtol=1e-13, atol=1e-14)
    S = orc.cma_scaledadd(pm, M, pc, C, pc, 2.4, 3.7)
    np.testing.assert_allclose(_dense(pc, S), 2.4 * _dense(pm, M) + 3.7 * _dense(pc, C), rtol=1e-13, atol=1e-14)
    # the tridiagonal solve of a diagonally dominant W3 column operator
    T = S.copy()
    T[:, :, 1] += 10.0
    b = rng.uniform(-1, 1, w3.undf)
    z = orc.cma_appinv(pc, T, b, w3.map, w3.ind)
    np.testing.assert_allclose(np.einsum("cij,cj->ci", _dense(pc, T), _column_values(w3, z)),
                               _column_values(w3, b), rtol=1e-12, atol=1e-13)
    # lumped inverse: reciprocal of the assembled diagonal
    lma = rng.uniform(0.5, 1.5, (nc * nl, 2, 2))
    pl = orc.cma_params(nl, 0, 1, 0, 1, custom=(1, 1, 0, 0))
    L = orc.cma_asm_m2v_lumped_inv(nc, nl, lma, pl, th.banded)
    np.testing.assert_allclose(L[:, :, 0], 1.0 / np.diagonal(_dense_from_lma(th, th, lma), axis1=1, axis2=2),
                               rtol=1e-14)
    # Dh diag(M2h)^-1 Dh^T on W3 columns
    Dh = rng.uniform(-1, 1, (nc * nl, 1, 4))
    M2h = rng.uniform(-1, 1, (nc * nl, 4, 4)) + 4 * np.eye(4)
    H = orc.cma_asm_diag_hmht(nc, nl, Dh, M2h, pm, w3.banded)
    want = np.einsum("ed,ed,ed->e", Dh[:, 0, :], Dh[:, 0, :], 1.0 / np.diagonal(M2h, axis1=1, axis2=2))
    np.testing.assert_allclose(H[:, :, 0].ravel(), want, rtol=1e-13)


def test_integer_band_limits_equal_the_reference_real_arithmetic():
    """The device computes j_minus = ceiling((alpha i - gamma_p) / beta) and j_plus = floor((alpha i
    + gamma_m) / beta) in integer arithmetic; the reference divides in r_solver reals.  They agree
    over every parameter set a lowest-order operator can have and rows up to 2^14."""
    for alpha in range(1, 5):
        for beta in range(1, 5):
            for g in range(0, 9):
                i = np.arange(1, 1 << 14, dtype=np.int64)
                num_m, num_p = alpha * i - g, alpha * i + g
                fm = np.ceil(num_m.astype(np.float32) / np.float32(beta)).astype(np.int64)
                fp = np.floor(num_p.astype(np.float32) / np.float32(beta)).astype(np.int64)
                np.testing.assert_array_equal(fm, -((-num_m) // beta))
                np.testing.assert_array_equal(fp, num_p // beta)


def test_tri_solve_known_answer_and_equals_appinv(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["tri_solve"]
    n = len(k["x"])
    y = orc.tri_solve(n, [1], k["x"], np.full(n, k["A0"]), np.full(n, k["AP"]), np.full(n, k["AM"]))
    np.testing.assert_allclose(y, k["answer"], rtol=0, atol=1e-14)
    # the same arithmetic as the CMA tridiagonal solve on bands (AM, A0, AP), bit for bit
    rng = np.random.default_rng(3)
    nc, nl = 4, 9
    w3 = _Space(orc, "w3", nc, nl)
    for r32 in (False, True):
        T = np.float32 if r32 else np.float64
        t0 = rng.uniform(3, 5, w3.undf).astype(T)
        tp, tm, x = (rng.uniform(-1, 1, w3.undf).astype(T) for _ in range(3))
        ys = orc.tri_solve(nl, w3.map, x, t0, tp, tm, r32=r32)
        p = orc.cma_params(nl, 1, 0, 1, 0, custom=(1, 1, 1, 1))
        bands = np.stack([tm, t0, tp], axis=-1).reshape((nc, nl, 3))
        yc = orc.cma_appinv(p, bands, x, w3.map, w3.ind, r32=r32)
        np.testing.assert_array_equal(ys, yc)


def _ref():
    path = ROOT / "oracle" / "_ref" / "libref.so"
    if not path.exists():
        pytest.skip("oracle/_ref not built (needs the reference tree and flang)")
    return ctypes.CDLL(str(path))


def test_cma_bitwise_equals_compiled_reference_kernels(orc):
    R = _ref()
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    I = ctypes.c_int
    rng = np.random.default_rng(4)
    nc, nl = 6, 7
    w3, th = _Space(orc, "w3", nc, nl), _Space(orc, "wtheta", nc, nl)
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    cmas = {}
    for to, fr in ((w3, th), (th, w3), (th, th), (w3, w3)):
        lma = f32(rng.uniform(-1, 1, (nc * nl, to.ndf, fr.ndf)))
        p = _params(orc, to, fr)
        A = orc.cma_asm(nc, nl, lma, p, to.banded, fr.banded, r32=True)
        Ar = np.full_like(A, np.nan)
        lf = f32(orc.flat_matrix(lma))  # kept alive across the call
        R.ref_cma_asm_r32(I(nc), I(nl), I(to.ndf), I(fr.ndf), P(lf), P(p), P(to.banded), P(fr.banded), P(Ar))
        np.testing.assert_array_equal(A, Ar)
        x = f32(rng.uniform(-1, 1, fr.undf))
        y0 = f32(rng.uniform(-1, 1, to.undf))  # the kernel accumulates onto lhs
        y = orc.cma_apply(p, A, x, y0, to.map, to.ind, fr.map, fr.ind, r32=True)
        yr = y0.copy()
        R.ref_cma_apply_r32(I(nc), P(p), P(A), P(x), P(yr), I(to.ndf), I(to.undf), P(to.map), P(to.ind), I(fr.ndf),
                            I(fr.undf), P(fr.map), P(fr.ind))
        np.testing.assert_array_equal(y, yr)
        cmas[(id(to), id(fr))] = (p, A)
    pa, A = cmas[(id(w3), id(th))]
    pb, B = cmas[(id(th), id(w3))]
    pm, M = cmas[(id(w3), id(w3))]
    pc = orc.cma_params(nl, 1, 0, 1, 0, custom=(1, 1, 1, 1))
    C0 = f32(rng.uniform(-1, 1, (nc, pc[5], pc[4])))
    C = orc.cma_mul(pa, A, pb, B, pc, C0, r32=True)
    Cr = C0.copy()
    R.ref_cma_mul_r32(I(nc), P(pa), P(A), P(pb), P(B), P(pc), P(Cr))
    np.testing.assert_array_equal(C, Cr)
    S = orc.cma_scaledadd(pm, M, pc, C, pc, 2.4, 3.7, r32=True)
    Sr = np.full_like(S, np.nan)
    R.ref_cma_scaledadd_r32(I(nc), P(pm), P(M), P(pc), P(C), P(pc), P(Sr), ctypes.c_float(2.4), ctypes.c_float(3.7))
    np.testing.assert_array_equal(S, Sr)
    S[:, :, 1] += np.float32(8.0)
    b = f32(rng.uniform(-1, 1, w3.undf))
    z = orc.cma_appinv(pc, S, b, w3.map, w3.ind, r32=True)
    zr = np.zeros_like(z)
    R.ref_cma_appinv_r32(I(nc), P(pc), P(S), P(b), P(zr), I(1), I(w3.undf), P(w3.map), P(w3.ind))
    np.testing.assert_array_equal(z, zr)
    Dh = f32(rng.uniform(-1, 1, (nc * nl, 1, 4)))
    M2h = f32(rng.uniform(-1, 1, (nc * nl, 4, 4)) + 4 * np.eye(4))
    H = orc.cma_asm_diag_hmht(nc, nl, Dh, M2h, pm, w3.banded, r32=True)
    Hr = np.full_like(H, np.nan)
    Df, Mf = f32(orc.flat_matrix(Dh)), f32(orc.flat_matrix(M2h))
    R.ref_cma_asm_diag_hmht_r32(I(nc), I(nl), I(1), I(4), P(Df), P(Mf), P(pm), P(w3.banded), P(Hr))
    np.testing.assert_array_equal(H, Hr)
    lma = f32(rng.uniform(-1, 1, (nc * nl, 2, 2)))
    lma[3, 0, 0] = -lma[3, 1, 1]  # a row whose lumped sum can vanish stays zero
    pl = orc.cma_params(nl, 0, 1, 0, 1, custom=(1, 1, 0, 0))
    L = orc.cma_asm_m2v_lumped_inv(nc, nl, lma, pl, th.banded, r32=True)
    Lr = np.full_like(L, np.nan)
    lf = f32(orc.flat_matrix(lma))
    R.ref_cma_asm_m2v_lumped_inv_r32(I(nc), I(nl), I(2), P(lf), P(pl), P(th.banded), P(Lr))
    np.testing.assert_array_equal(L, Lr)
