"""On-device assembly of lowest-order W0 mass matrices from Cartesian coordinate fields
(compute_mass_matrix_w_scalar_code), against the oracle and the reference's unit-test answer.

Bars: every entry equal to the oracle's bit for bit (both sum the same products in the same
order without contraction, from the same Gauss rule and basis tables); the reference's known
answer within its own tolerance; a CG solve with the assembled matrix recovers x = 1 from
b = M 1."""
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
