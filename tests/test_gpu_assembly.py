"""On-device assembly of lowest-order mass matrices (W0: compute_mass_matrix_w_scalar_code, W2:
compute_mass_matrix_w2_code, W3: compute_mass_matrix_w3_code) and of the weak divergence W2 -> W3
(compute_div_operator_code) from Cartesian coordinate fields, against the oracle and the
reference's unit-test answers.

Bars: every entry equal to the oracle's bit for bit (both sum the same products in the same
order without contraction, from the same Gauss rule and basis tables); the reference's known
answers within their own tolerance; a CG solve with the assembled W0 matrix recovers x = 1 from
b = M 1; the Helmholtz operator M3 + tau D diag(M2)^-1 D^T built from assembled operators equals
its numpy restatement from the oracle's matrices, and GMRES recovers a known solution."""
import json

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _box_mesh(nx, ny, nl):
    """A structured nx x ny box of cell columns (not periodic): vertex columns numbered
    row-major, each nl+1 dofs high (bottom dof then top dof of a layer = +1), cells in four
    checkerboard colours."""
    H = nl + 1
    vid = lambda i, j: j * (nx + 1) + i
    cells, colours = [], [[] for _ in range(4)]
    for j in range(ny):
        for i in range(nx):
            v = [vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)]  # SW, SE, NE, NW
            cells.append([x * H + 1 for x in v] + [x * H + 2 for x in v])
            colours[(i % 2) + 2 * (j % 2)].append(len(cells))
    dm = np.array(cells, np.int32)
    npc = np.array([len(c) for c in colours], np.int32)
    cic = np.zeros((4, npc.max()), np.int32)
    for c in range(4):
        cic[c, : npc[c]] = colours[c]
    return dm, npc, cic, (nx + 1) * (ny + 1) * H


def _chi_of_box(nx, ny, nl, rng):
    H = nl + 1
    chi = np.zeros((3, (nx + 1) * (ny + 1) * H))
    z = np.cumsum(np.concatenate([[0.0], rng.uniform(0.5, 1.5, nl)]))  # stretched levels
    for j in range(ny + 1):
        for i in range(nx + 1):
            base = (j * (nx + 1) + i) * H
            xy = np.array([i * 1.0, j * 0.8]) + rng.uniform(-0.15, 0.15, 2)  # perturbed grid
            for k in range(H):
                chi[:, base + k] = (xy[0], xy[1], z[k] + 0.05 * xy[0])
    return chi


def test_assembled_mass_matrix_bitwise_equals_oracle_and_solves(hip, orc):
    nx, ny, nl = 6, 5, 4
    dm, npc, cic, undf = _box_mesh(nx, ny, nl)
    fs = hip.FunctionSpace.from_dofmap(nl, undf, dm, npc, cic)
    chi = _chi_of_box(nx, ny, nl, np.random.default_rng(5))
    cv = [hip.FieldVector(fs).import_field(c) for c in chi]
    for nqp in (2, 3, 4):
        op = hip.LmaOperator(fs).assemble_mass_w0(*cv, nqp=nqp)
        mm = orc.mass_matrix_w0(nx * ny, nl, dm, chi[0], chi[1], chi[2], nqp=nqp)
        np.testing.assert_array_equal(op.download(), mm)
    # M x = M 1 has the solution x = 1 (M is SPD on this non-degenerate mesh)
    op = hip.LmaOperator(fs).assemble_mass_w0(*cv)
    one = hip.FieldVector(fs)
    one.set_scalar(1.0)
    b = hip.FieldVector(fs)
    op.apply(one, b)
    x = hip.FieldVector(fs)
    x.set_scalar(0.0)
    prec = hip.DiagonalPreconditioner.from_operator(op)
    res = hip.ConjugateGradient(op, prec, 1e-12, 0.0, 500, check_every=4).apply(x, b)
    assert res.status == 0
    np.testing.assert_allclose(x.export_field(), 1.0, rtol=0, atol=1e-9)


def test_assembled_mass_matrix_known_answer(hip, orc):
    """The 6000 x 1000 x 2000 box cell of compute_mass_matrix_kernel_w0_mod_test.pf."""
    k = json.loads((GOLDEN / "kat.json").read_text())["mass_matrix_w0"]
    nl, dx, dy, dz = k["nlayers"], k["dx"], k["dy"], k["dz"]
    dm, npc, cic, undf = _box_mesh(1, 1, nl)
    fs = hip.FunctionSpace.from_dofmap(nl, undf, dm, npc, cic)
    chi = np.zeros((3, undf))
    H = nl + 1
    for j in range(2):
        for i in range(2):
            for kk in range(H):
                chi[:, (j * 2 + i) * H + kk] = (i * dx, j * dy, kk * dz)
    cv = [hip.FieldVector(fs).import_field(c) for c in chi]
    mm = hip.LmaOperator(fs).assemble_mass_w0(*cv).download()
    ans = np.array(k["answer_column_major"]).reshape((8, 8), order="F")
    vol = dx * dy * dz
    for layer in range(nl):
        np.testing.assert_allclose(mm[layer] / vol, ans / vol, rtol=0, atol=k["tol"])
    with pytest.raises(hip.LfricError):  # FP32 operators are converted after assembly
        hip.LmaOperator(fs, precision=hip.R32).assemble_mass_w0(*cv)


def _box_w2_w3(nx, ny, nl):
    """Lowest-order W2 and W3 dofmaps of the box of _box_mesh (same cells, same colouring):
    x-normal face columns (nl dofs each), then y-normal ones, then one column of horizontal faces
    per cell (nl + 1 dofs: the top face of layer k is the bottom face of layer k + 1).  Local
    order W, S, E, N, B, T."""
    xcol = lambda i, j: (j * (nx + 1) + i) * nl
    ny0 = (nx + 1) * ny * nl
    ycol = lambda i, j: ny0 + (j * nx + i) * nl
    nz0 = ny0 + nx * (ny + 1) * nl
    zcol = lambda i, j: nz0 + (j * nx + i) * (nl + 1)
    w2, w3 = [], []
    for j in range(ny):
        for i in range(nx):
            b = zcol(i, j)
            w2.append([xcol(i, j) + 1, ycol(i, j) + 1, xcol(i + 1, j) + 1, ycol(i, j + 1) + 1, b + 1, b + 2])
            w3.append([(j * nx + i) * nl + 1])
    return np.array(w2, np.int32), nz0 + nx * ny * (nl + 1), np.array(w3, np.int32), nx * ny * nl


def _box_spaces(hip, nx, ny, nl):
    dm0, npc, cic, undf0 = _box_mesh(nx, ny, nl)
    dm2, undf2, dm3, undf3 = _box_w2_w3(nx, ny, nl)
    w0 = hip.FunctionSpace.from_dofmap(nl, undf0, dm0, npc, cic)
    w2 = hip.FunctionSpace.from_dofmap(nl, undf2, dm2, npc, cic)
    w3 = hip.FunctionSpace.from_dofmap(nl, undf3, dm3, npc, cic)
    return (w0, dm0), (w2, dm2), (w3, dm3)


def test_assembled_w2_w3_div_bitwise_equal_oracle(hip, orc):
    nx, ny, nl = 5, 4, 3
    (w0, dm0), (w2, _), (w3, _) = _box_spaces(hip, nx, ny, nl)
    chi = _chi_of_box(nx, ny, nl, np.random.default_rng(7))
    cv = [hip.FieldVector(w0).import_field(c) for c in chi]
    nc = nx * ny
    for nqp in (2, 3, 4):
        print("assemble nqp", nqp, flush=True)
        m2 = hip.LmaOperator(w2).assemble_mass(*cv, nqp=nqp).download()
        np.testing.assert_array_equal(m2, orc.mass_matrix_w2(nc, nl, dm0, *chi, nqp=nqp))
        for rehab in (True, False):
            m3 = hip.LmaOperator(w3).assemble_mass(*cv, nqp=nqp, rehabilitate=rehab).download()
            np.testing.assert_array_equal(m3, orc.mass_matrix_w3(nc, nl, dm0, *chi, nqp=nqp, rehabilitate=rehab))
            d = hip.LmaOperator(w3, w2).assemble_div(*cv, nqp=nqp, rehabilitate=rehab).download()
            np.testing.assert_array_equal(d, orc.div_operator(nc, nl, dm0, *chi, nqp=nqp, rehabilitate=rehab))
    # the W0 path of the general entry point is the W0 assembly
    m0 = hip.LmaOperator(w0).assemble_mass(*cv).download()
    np.testing.assert_array_equal(m0, orc.mass_matrix_w0(nc, nl, dm0, *chi, nqp=3))


def test_assembled_w2_w3_div_known_answers_and_refusals(hip):
    kat = json.loads((GOLDEN / "kat.json").read_text())
    k2, k3, kd = kat["mass_matrix_w2"], kat["mass_matrix_w3"], kat["div_operator"]
    nl, dx, dy, dz = k2["nlayers"], k2["dx"], k2["dy"], k2["dz"]
    (w0, dm0), (w2, _), (w3, _) = _box_spaces(hip, 1, 1, nl)
    chi = np.zeros((3, w0.undf))
    H = nl + 1
    for j in range(2):
        for i in range(2):
            for kk in range(H):
                chi[:, (j * 2 + i) * H + kk] = (i * dx, j * dy, kk * dz)
    cv = [hip.FieldVector(w0).import_field(c) for c in chi]
    m2 = hip.LmaOperator(w2).assemble_mass(*cv).download()
    ans = np.array(k2["answer_column_major"]).reshape((6, 6), order="F")
    norm = (dx + dy + dz) / 3
    m3 = hip.LmaOperator(w3).assemble_mass(*cv).download()
    d = hip.LmaOperator(w3, w2).assemble_div(*cv).download()
    for layer in range(nl):
        np.testing.assert_allclose(m2[layer] * norm, ans * norm, rtol=0, atol=k2["tol"])
        np.testing.assert_allclose(m3[layer, 0, 0] / (dx * dy * dz), k3["answer_over_volume"], rtol=0, atol=k3["tol"])
        np.testing.assert_allclose(d[layer, 0], kd["answer"], rtol=0, atol=kd["tol"])
    for bad in (lambda: hip.LmaOperator(w2, precision=hip.R32).assemble_mass(*cv),  # FP32: convert afterwards
                lambda: hip.LmaOperator(w2).assemble_div(*cv),                      # a divergence maps W2 to W3
                lambda: hip.LmaOperator(w3, w2).assemble_mass(*cv),                 # a mass matrix is square
                lambda: hip.LmaOperator(w2).assemble_mass(*cv, nqp=5),
                lambda: hip.LmaOperator(w2).assemble_mass(cv[0], cv[1], hip.FieldVector(w2))):
        with pytest.raises(hip.LfricError):
            bad()


def test_helmholtz_from_assembled_operators(hip, orc):
    """H = M3 + tau D diag(M2)^-1 D^T with every operator assembled on the device equals the
    numpy restatement built from the oracle's element matrices, and FP64 GMRES recovers a known
    solution of H x = b."""
    nx, ny, nl, tau = 6, 5, 4, 0.7
    (w0, dm0), (w2, dm2), (w3, dm3) = _box_spaces(hip, nx, ny, nl)
    chi = _chi_of_box(nx, ny, nl, np.random.default_rng(9))
    cv = [hip.FieldVector(w0).import_field(c) for c in chi]
    M2 = hip.LmaOperator(w2).assemble_mass(*cv)
    M3 = hip.LmaOperator(w3).assemble_mass(*cv)
    D = hip.LmaOperator(w3, w2).assemble_div(*cv)
    H = hip.LinearOperator.helmholtz(M3, D, M2.mm_diagonal(hip.FieldVector(w2)), tau)
    # dense restatement from the oracle's element matrices and the dofmaps
    nc = nx * ny
    m2 = orc.mass_matrix_w2(nc, nl, dm0, *chi)
    m3 = orc.mass_matrix_w3(nc, nl, dm0, *chi)
    dd = orc.div_operator(nc, nl, dm0, *chi)
    n2, n3 = w2.undf, w3.undf
    diag2 = np.zeros(n2)
    Dg = np.zeros((n3, n2))
    M3g = np.zeros((n3, n3))
    for c in range(nc):
        for k in range(nl):
            ik = c * nl + k
            r2 = dm2[c] - 1 + k
            r3 = dm3[c, 0] - 1 + k
            diag2[r2] += np.diag(m2[ik])
            M3g[r3, r3] += m3[ik, 0, 0]
            Dg[r3, r2] += dd[ik, 0]
    Hg = M3g + tau * Dg @ np.diag(1.0 / diag2) @ Dg.T
    x = np.random.default_rng(10).uniform(-1, 1, n3)
    y = hip.FieldVector(w3)
    H.apply(hip.FieldVector(w3).import_field(x), y)
    np.testing.assert_allclose(y.export_field(), Hg @ x, rtol=1e-12, atol=1e-12 * np.abs(Hg @ x).max())
    b = hip.FieldVector(w3).import_field(Hg @ x)
    xs = hip.FieldVector(w3)
    res = hip.GMRES(H, hip.NullPreconditioner(), 30, 1e-12, 0.0, 200).apply(xs, b)
    assert res.status == 0
    np.testing.assert_allclose(xs.export_field(), x, rtol=0, atol=1e-8)


def test_assembled_wtheta_mass_bitwise_equals_oracle_and_known_answer(hip, orc):
    """The Wtheta mass matrix (compute_mass_matrix_w_scalar_code with the Wtheta basis) on the
    box: bit-equal to the oracle for 2-4 points per direction; 1/3 and 1/6 of the volume on the
    reference unit test's box cell."""
    nx, ny, nl = 5, 4, 3
    dm0, npc, cic, undf0 = _box_mesh(nx, ny, nl)
    nc = nx * ny
    dmt = np.array([[c * (nl + 1) + 1, c * (nl + 1) + 2] for c in range(nc)], np.int32)
    w0 = hip.FunctionSpace.from_dofmap(nl, undf0, dm0, npc, cic)
    wt = hip.FunctionSpace.from_dofmap(nl, nc * (nl + 1), dmt, npc, cic)
    chi = _chi_of_box(nx, ny, nl, np.random.default_rng(15))
    cv = [hip.FieldVector(w0).import_field(c) for c in chi]
    for nqp in (2, 3, 4):
        m = hip.LmaOperator(wt).assemble_mass(*cv, nqp=nqp).download()
        np.testing.assert_array_equal(m, orc.mass_matrix_wtheta(nc, nl, dm0, *chi, nqp=nqp))
    k = json.loads((GOLDEN / "kat.json").read_text())["mass_matrix_w3"]  # the unit tests' box
    dx, dy, dz = k["dx"], k["dy"], k["dz"]
    dm1, npc1, cic1, undf1 = _box_mesh(1, 1, nl)
    w01 = hip.FunctionSpace.from_dofmap(nl, undf1, dm1, npc1, cic1)
    wt1 = hip.FunctionSpace.from_dofmap(nl, nl + 1, np.array([[1, 2]], np.int32), npc1, cic1)
    c1 = np.zeros((3, undf1))
    for j in range(2):
        for i in range(2):
            for kk in range(nl + 1):
                c1[:, (j * 2 + i) * (nl + 1) + kk] = (i * dx, j * dy, kk * dz)
    m = hip.LmaOperator(wt1).assemble_mass(*[hip.FieldVector(w01).import_field(c) for c in c1]).download()
    vol = dx * dy * dz
    for ik in range(nl):
        np.testing.assert_allclose(m[ik] / vol, [[1 / 3, 1 / 6], [1 / 6, 1 / 3]], rtol=0, atol=1e-6)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_cubed_sphere_coordinates_and_mass_matrices_bitwise_equal_oracle(hip, orc, kind):
    """Coordinate fields of the generated cubed sphere (mesh generator vertices, extrusion,
    spherical shells) equal the oracle's bit for bit; the W0 and W3 mass matrices assembled from
    them on the device equal the oracle's; M x = M 1 solves to x = 1."""
    n, nl = 6, 8
    fs = hip.FunctionSpace.cubedsphere(n, nl)
    m = orc.Mesh(n, nl)
    np.testing.assert_array_equal(hip.extrusion_eta(kind, nl), orc.eta(kind, nl))
    cv = fs.coordinates(kind, 80000.0, 6371229.0)
    chi = orc.chi_w0(m, kind, 80000.0, 6371229.0)
    for a, b in zip(cv, chi):
        np.testing.assert_array_equal(a.export_field(), b)
    op = hip.LmaOperator(fs).assemble_mass_w0(*cv, nqp=3)
    np.testing.assert_array_equal(op.download(), orc.mass_matrix_w0(m.ncells, nl, m.dofmap, *chi, nqp=3))
    w3 = fs.sibling(hip.W3, 0)
    m3 = hip.LmaOperator(w3).assemble_mass(*cv, nqp=3)
    np.testing.assert_array_equal(m3.download().ravel(), orc.mass_matrix_w3(m.ncells, nl, m.dofmap, *chi, nqp=3).ravel())
    # the cells' volumes add up to the spherical shell's, within the trilinear cells' chord error
    R, top = 6371229.0, 80000.0
    shell = 4.0 / 3.0 * np.pi * ((R + top) ** 3 - R ** 3)
    assert abs(m3.download().sum() / shell - 1.0) < 0.05
    one = hip.FieldVector(fs)
    one.set_scalar(1.0)
    b = hip.FieldVector(fs)
    op.apply(one, b)
    x = hip.FieldVector(fs)
    x.set_scalar(0.0)
    prec = hip.DiagonalPreconditioner.from_operator(op)
    res = hip.ConjugateGradient(op, prec, 1e-12, 0.0, 2000, check_every=8).apply(x, b)
    assert res.status == 0
    np.testing.assert_allclose(x.export_field(), 1.0, rtol=0, atol=1e-8)


def test_c96x70_quadratic_mass_matrix_assembles_and_solves(hip):
    """The W0 mass matrix of the C96 cubed sphere with 70 quadratically stretched levels (80 km
    lid), assembled on the device and solved with mixed-precision CG (FP32 inner, FP64
    refinement) for b = M 1."""
    fs = hip.FunctionSpace.cubedsphere(96, 70)
    cv = fs.coordinates(1, 80000.0, 6371229.0)
    op64 = hip.LmaOperator(fs).assemble_mass_w0(*cv, nqp=3)
    op32 = op64.convert(hip.R32)
    one = hip.FieldVector(fs)
    one.set_scalar(1.0)
    b = hip.FieldVector(fs)
    op64.apply(one, b)
    x = hip.FieldVector(fs)
    x.set_scalar(0.0)
    prec32 = hip.DiagonalPreconditioner.from_operator(op32)
    res = hip.MixedPrecisionCG(op64, op32, prec32, 1e-10, 0.0, 20000, inner_r_tol=1e-4).apply(x, b)
    print(f"C96x70 quadratic W0 mass solve: {res.iterations} inner, {res.outer_iterations} outer, "
          f"true rel res {res.true_rel_res:.2e}", flush=True)
    assert res.status == 0 and res.true_rel_res <= 1e-10
    np.testing.assert_allclose(x.export_field(), 1.0, rtol=0, atol=1e-6)
