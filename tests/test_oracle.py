"""The oracle pinned against the reference's own fixtures, known-answer tests, the compiled
reference kernel (oracle/_ref) and closed-form / brute-force checks (CPU only)."""
import ctypes
import hashlib
import json

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from conftest import GOLDEN, ROOT


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


# ---- topology: the committed UGRID meshes of the reference --------------------------------
@pytest.mark.parametrize("n", [4, 8, 12, 16, 24, 32])
def test_cubedsphere_topology_matches_ugrid_fixtures(orc, n):
    cn, vc = orc.cubedsphere_topology(n)
    dig = json.loads((GOLDEN / "mesh_digests.json").read_text())
    assert _digest(vc) == dig[f"C{n}_face_nodes"]
    assert _digest(cn) == dig[f"C{n}_face_links"]
    if n <= 12:
        fx = np.load(GOLDEN / "mesh_fixtures.npz")
        np.testing.assert_array_equal(vc, fx[f"C{n}_face_nodes"])
        np.testing.assert_array_equal(cn, fx[f"C{n}_face_links"])


def test_c3_counts_match_mesh_generator_summary(orc):
    kat = json.loads((GOLDEN / "kat.json").read_text())["c3_summary"]
    cn, vc = orc.cubedsphere_topology(3)
    lid, nv = orc.local_vert_lids(vc)
    foc, eoc, nf, ne = orc.mesh_connectivity(cn, lid)
    assert (nv, ne, cn.shape[0]) == (kat["Nodes"], kat["Edges"], kat["Faces"])


def test_topology_invariants(orc):
    for n in (2, 5, 12):
        cn, vc = orc.cubedsphere_topology(n)
        nc = 6 * n * n
        # adjacency symmetric, every cell has 4 distinct neighbours
        for c in range(nc):
            assert len(set(cn[c])) == 4
            for d in range(4):
                assert c + 1 in cn[cn[c, d] - 1]
        lid, nv = orc.local_vert_lids(vc)
        assert nv == nc + 2
        counts = np.bincount(lid.ravel())[1:]
        assert sorted(set(counts)) == [3, 4] and (counts == 3).sum() == 8


# ---- connectivity and dofmaps: canned bi-periodic mesh fixtures ---------------------------
def test_connectivity_and_dofmaps_match_unit_test_fixtures(orc):
    c = json.loads((GOLDEN / "canned_mesh.json").read_text())
    cn = np.array(c["cell_next"], np.int32)
    vc = np.array(c["vert_on_cell"], np.int32)
    foc, eoc, nf, ne = orc.mesh_connectivity(cn, vc)
    np.testing.assert_array_equal(foc, c["face_on_cell"])
    np.testing.assert_array_equal(eoc, c["edge_on_cell"])
    nv = int(vc.max())
    for space, key in ((orc.W0, "fs0k0_dofmap"), (orc.W1, "fs1k0_dofmap"), (orc.W2, "fs2k0_dofmap"),
                       (orc.W3, "fs3k0_dofmap")):
        dm, undf = orc.dofmap(space, 0, c["nlayers"], vc, foc, eoc, nv, ne)
        np.testing.assert_array_equal(dm, c[key])


def test_w1_fixture_pins_the_edge_branch(orc):
    """fs1k0_dofmap holds only edge dofs: the bottom horizontal edges get nlayers+1 ids with the
    top edge one above, the vertical edges nlayers ids, numbered cell by cell with h-edges first.
    W0 order 1 takes its edge dofs from this same branch, so a swap of the h-/v-edge loops or a
    wrong top offset in the oracle would break the fixture equality above; check that it does."""
    c = json.loads((GOLDEN / "canned_mesh.json").read_text())
    fx = np.array(c["fs1k0_dofmap"])
    nl = c["nlayers"]
    np.testing.assert_array_equal(fx[:, 8:12], fx[:, 0:4] + 1)
    nh, nvv = np.unique(fx[:, 0:4]).size, np.unique(fx[:, 4:8]).size
    undf = nh * (nl + 1) + nvv * nl
    starts = np.unique(np.concatenate([fx[:, 0:4].ravel(), fx[:, 4:8].ravel()]))
    lengths = dict(zip(starts, np.diff(np.append(starts, undf + 1))))
    assert {lengths[v] for v in fx[:, 0:4].ravel()} == {nl + 1}
    assert {lengths[v] for v in fx[:, 4:8].ravel()} == {nl}
    # cell 1 numbers its four bottom h-edges before its vertical edges
    assert fx[0, :4].max() < fx[0, 4:8].min()


# W0 order 1: 27 local dofs in entity order volume (0), side faces (1-4), bottom and top face
# (5, 6), bottom h-edges (7-10), vertical edges (11-14), top h-edges (15-18), bottom vertices
# (19-22), top vertices (23-26)
_W0K1_KIND = ["vol"] + ["vface"] * 4 + ["hface_b", "hface_t"] + ["hedge_b"] * 4 + ["vedge"] * 4 + \
             ["hedge_t"] * 4 + ["vert_b"] * 4 + ["vert_t"] * 4
_W0K1_TOP_OF = {6: 5, **{15 + i: 7 + i for i in range(4)}, **{23 + i: 19 + i for i in range(4)}}


def _w0k1_entity_violations(dm, undf, nl):
    """Entity structure of a W0 order-1 dofmap on a cubed sphere, independent of the numbering
    rule: every column is nl+1 long for vertices, horizontal edges and horizontal faces and nl
    long otherwise; each top entity is its bottom partner plus one; each column is reached from
    one kind of local position only, by as many cells as share that entity (volume 1, side face 2,
    horizontal face 1, h-edge 2, vertical edge and vertex 3 or 4).  Returns the failed checks."""
    bad = []
    for t, b in _W0K1_TOP_OF.items():
        if not np.array_equal(dm[:, t], dm[:, b] + 1):
            bad.append(f"top {t} != bottom {b} + 1")
    base = [p for p in range(27) if p not in _W0K1_TOP_OF]
    starts = np.unique(dm[:, base])
    length = dict(zip(starts.tolist(), np.diff(np.append(starts, undf + 1)).tolist()))
    want_len = {"vol": nl, "vface": nl, "vedge": nl, "hface_b": nl + 1, "hedge_b": nl + 1, "vert_b": nl + 1}
    kind_of, cells_of = {}, {}
    for p in base:
        k = _W0K1_KIND[p]
        for cell, col in enumerate(dm[:, p].tolist()):
            if kind_of.setdefault(col, k) != k:
                bad.append(f"column {col} reached as {kind_of[col]} and {k}")
            cells_of.setdefault(col, set()).add(cell)
            if length.get(col) != want_len[k]:
                bad.append(f"{k} column {col} has length {length.get(col)}")
    share = {"vol": {1}, "vface": {2}, "hface_b": {1}, "hedge_b": {2}, "vedge": {3, 4}, "vert_b": {3, 4}}
    for col, k in kind_of.items():
        if len(cells_of[col]) not in share[k]:
            bad.append(f"{k} column {col} shared by {len(cells_of[col])} cells")
    if sum(length.values()) != undf:
        bad.append("columns do not tile 1..undf")
    return bad


@pytest.mark.parametrize("n,nl", [(3, 4), (6, 5)])
def test_w0_order1_dofmap_entity_sharing(orc, n, nl):
    m = orc.Mesh(n, nl, orc.W0, 1)
    assert _w0k1_entity_violations(m.dofmap, m.undf, nl) == []
    nv = 6 * n * n + 2
    # vertical edges sit on the 2D vertices: 8 of their columns are shared by 3 cells
    vcols = np.unique(m.dofmap[:, 11:15])
    assert vcols.size == nv and sum(len(set(np.where(m.dofmap[:, 11:15] == v)[0])) == 3 for v in vcols) == 8
    # planted mistakes in a copy of the oracle's dofmap: h-edges and v-edges swapped, and top
    # h-edges two above their bottom partners; both must be caught
    swapped = m.dofmap.copy()
    swapped[:, 7:11], swapped[:, 11:15] = m.dofmap[:, 11:15], m.dofmap[:, 7:11]
    assert _w0k1_entity_violations(swapped, m.undf, nl)
    shifted = m.dofmap.copy()
    shifted[:, 15:19] += 1
    assert _w0k1_entity_violations(shifted, m.undf, nl)


@pytest.mark.parametrize("n,nl", [(4, 3), (12, 20)])
def test_undf_closed_forms_and_column_contiguity(orc, n, nl):
    nc, ne, nv = 6 * n * n, 12 * n * n, 6 * n * n + 2
    m0 = orc.Mesh(n, nl, orc.W0, 0)
    assert m0.undf == nv * (nl + 1)
    m1 = orc.Mesh(n, nl, orc.W0, 1)
    assert m1.undf == (nc + ne + nv) * (2 * nl + 1)
    # W0 order 1: vertex/h-edge/h-face columns nl+1, volume/side faces/v-edges nl
    assert m1.undf == nv * (nl + 1) + ne * (nl + 1) + nc * (nl + 1) + nc * nl + ne * nl + nv * nl
    assert orc.Mesh(n, nl, orc.W2, 0).undf == ne * nl + nc * (nl + 1)
    assert orc.Mesh(n, nl, orc.W3, 0).undf == nc * nl
    # every W0 vertex column is shared by 4 cells (3 at the 8 cube corners)
    counts = np.bincount(m0.dofmap[:, :4].ravel())
    counts = counts[counts > 0]
    assert (counts == 3).sum() == 8 and set(counts) == {3, 4}
    np.testing.assert_array_equal(m0.dofmap[:, 4:], m0.dofmap[:, :4] + 1)


# ---- colouring ------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [4, 12, 24])
def test_colouring_closed_form(orc, n):
    m = orc.Mesh(n, 2)
    assert list(m.ncells_per_colour) == [6 * n * n // 4] * 4
    assert orc.colour_validate(m.cell_next, m.colour_map) == 1
    # brute force: no two cells of one colour share a vertex
    for col in range(1, 5):
        cells = np.where(m.colour_map == col)[0]
        verts = m.vert_on_cell[cells].ravel()
        assert len(np.unique(verts)) == verts.size
    # the validator rejects a planted clash
    bad = m.colour_map.copy()
    bad[0] = bad[m.cell_next[0, 2] - 1]
    assert orc.colour_validate(m.cell_next, bad) == 0


@pytest.mark.parametrize("n", [3, 5, 7])
def test_odd_colouring_falls_back_to_greedy(orc, n):
    """Odd panels: the closed form does not validate, so set_colours takes the generic greedy
    colouring.  Checked independently from the vertex lists: each cell holds the lowest colour
    not held by a lower-numbered cell sharing a vertex with it, and no colour class shares a
    vertex."""
    m = orc.Mesh(n, 2)
    cm0 = np.zeros(m.ncells, np.int32)
    assert orc.lib().orc_colour_cubedsphere(n, orc._p(np.ascontiguousarray(m.cell_next, np.int32)),
                                            orc._p(cm0)) == 1
    cells_of_vert = {}
    for c in range(m.ncells):
        for v in m.vert_on_cell[c]:
            cells_of_vert.setdefault(v, []).append(c)
    for c in range(m.ncells):
        earlier = {m.colour_map[o] for v in m.vert_on_cell[c] for o in cells_of_vert[v] if o < c}
        assert m.colour_map[c] == min(set(range(1, 51)) - earlier)
    for col in range(1, m.ncolours + 1):
        verts = m.vert_on_cell[m.colour_map == col].ravel()
        assert len(np.unique(verts)) == verts.size
    assert m.ncells_per_colour.sum() == m.ncells and m.cells_in_colour.shape == (m.ncolours,
                                                                                  m.ncells_per_colour.max())


# ---- generator ------------------------------------------------------------------------------
def test_own_log_exp_against_libm(orc):
    L = orc.lib()
    xs = np.concatenate([np.linspace(1e-6, 1.0, 2001), [1e-300, 0.5, 0.999999, 1 - 2**-53]])
    for x in xs:
        assert abs(L.orc_log(x) - np.log(x)) <= 4 * np.spacing(abs(np.log(x))) + 1e-300
    for x in np.linspace(-9.3, 9.3, 2001):
        assert abs(L.orc_exp(x) - np.exp(x)) <= 4 * np.spacing(np.exp(x))


def test_generator_moments(orc):
    L = orc.lib()
    z = np.array([L.orc_rng_normal(0, 1, i) for i in range(20000)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.02
    u = orc.gen_uniform(20000, 0, 3)
    assert u.min() > -1 and u.max() < 1 and abs(u.mean()) < 0.02


def test_spd_elements_are_symmetric_positive_definite(orc):
    A = orc.gen_spd(8, 50, 0)
    assert np.array_equal(A, A.transpose(0, 2, 1))
    assert np.linalg.eigvalsh(A).min() > 0.0999
    A27 = orc.gen_spd(27, 4, 1, scaled=True)
    assert np.array_equal(A27, A27.transpose(0, 2, 1))


# ---- LMA apply -------------------------------------------------------------------------------
class _OneCell:
    """The one-cell, one-layer mesh of matrix_vector_mod_test.pf."""

    def __init__(self, dofmap, undf):
        self.ncells, self.nlayers = 1, 1
        self.dofmap = np.array([dofmap], np.int32)
        self.ndf, self.undf = len(dofmap), undf
        self.ncells_per_colour = np.array([1, 0, 0, 0], np.int32)
        self.cells_in_colour = np.array([[1], [0], [0], [0]], np.int32)


def test_lma_apply_known_answer(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["matrix_vector"]
    to, fr = _OneCell(k["map1"], 4), _OneCell(k["map2"], 2)
    A = np.array(k["d"])[None, :, :]
    y = orc.lma_apply(fr, A, np.array(k["x"]), to_mesh=to)
    np.testing.assert_allclose(y, k["answer"], rtol=0, atol=k["tol"])
    y32 = orc.lma_apply(fr, A, np.array(k["x"]), to_mesh=to, r32=True)
    np.testing.assert_allclose(y32, k["answer"], rtol=0, atol=1e-6)


@pytest.mark.parametrize("n,nl,order", [(4, 3, 0), (4, 3, 1), (12, 20, 0)])
def test_lma_apply_equals_assembled_matrix(orc, n, nl, order):
    m = orc.Mesh(n, nl, orc.W0, order)
    rng = np.random.default_rng(1)
    A = rng.standard_normal((m.ncell_3d, m.ndf, m.ndf))  # non-symmetric on purpose
    x = rng.standard_normal(m.undf)
    y = orc.lma_apply(m, A, x)
    S = orc.assembled_csr(m, A)
    bound = 1e-13 * (abs(S) @ np.abs(x))
    assert np.all(np.abs(y - S @ x) <= bound)
    # a transposed element matrix is caught
    yt = orc.lma_apply(m, A.transpose(0, 2, 1), x)
    assert np.abs(yt - y).max() > 1e-3


def test_mm_diagonal_is_assembled_diagonal(orc):
    m = orc.Mesh(4, 3, orc.W0, 1)
    A = orc.gen_spd(m.ndf, m.ncell_3d, 3)
    np.testing.assert_allclose(orc.mm_diagonal(m, A), orc.assembled_csr(m, A).diagonal(), rtol=1e-14)


def _ref_lib():
    p = ROOT / "oracle" / "_ref" / "libref.so"
    if not p.exists():
        pytest.skip("oracle/_ref not built (needs the reference tree and flang)")
    return ctypes.CDLL(str(p))


def test_oracle_apply_bitwise_equals_compiled_reference_kernel(orc):
    R = _ref_lib()
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    I = ctypes.c_int
    for order in (0, 1):
        m = orc.Mesh(4, 5, orc.W0, order)
        A = orc.gen_spd(m.ndf, m.ncell_3d, 7, scaled=True)
        x = orc.gen_uniform(m.undf, 0, 3)
        y = orc.lma_apply(m, A, x)
        yr = np.zeros(m.undf)
        Af = orc.flat_matrix(A)
        R.ref_lma_apply_r64(I(m.ncells), I(m.nlayers), I(4), P(m.ncells_per_colour), P(m.cells_in_colour),
                            I(m.cells_in_colour.shape[1]), I(m.ndf), I(m.undf), P(m.dofmap), I(m.ndf), I(m.undf),
                            P(m.dofmap), P(Af), P(x), P(yr))
        np.testing.assert_array_equal(y, yr)
        d = np.zeros(m.undf)
        R.ref_mm_diagonal_r64(I(m.ncells), I(m.nlayers), I(4), P(m.ncells_per_colour), P(m.cells_in_colour),
                              I(m.cells_in_colour.shape[1]), I(m.ndf), I(m.undf), P(m.dofmap), P(Af), P(d))
        np.testing.assert_allclose(d, orc.mm_diagonal(m, A), rtol=1e-15)


def test_transpose_and_dg_bitwise_equal_compiled_reference_kernels(orc):
    R = _ref_lib()
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    I = ctypes.c_int
    H = orc.Helmholtz(4, 3)
    w2, w3 = H.w2, H.w3
    x3 = orc.gen_uniform(w3.undf, 0, 3)
    y2 = orc.transpose_apply(w2, w3, H.D, x3)                # D^T: W3 -> W2, D is (c3, 1, 6)
    yr = np.zeros(w2.undf)
    R.ref_transpose_apply_r64(I(w2.ncells), I(w2.nlayers), I(4), P(w2.ncells_per_colour),
                              P(w2.cells_in_colour), I(w2.cells_in_colour.shape[1]), I(w2.ndf),
                              I(w2.undf), P(w2.dofmap), I(w3.ndf), I(w3.undf), P(w3.dofmap),
                              P(orc.flat_matrix(H.D)), P(x3), P(yr))
    np.testing.assert_array_equal(y2, yr)
    x2 = orc.gen_uniform(w2.undf, 1, 3)
    y3 = orc.dg_apply(w3, w2, H.D, x2)                       # D: W2 -> W3
    yr3 = np.zeros(w3.undf)
    R.ref_dg_apply_r64(I(w3.ncells), I(w3.nlayers), I(w3.ndf), I(w3.undf), P(w3.dofmap), I(w2.ndf),
                       I(w2.undf), P(w2.dofmap), P(orc.flat_matrix(H.D)), P(x2), P(yr3))
    np.testing.assert_array_equal(y3, yr3)


def test_transpose_known_answer(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["transpose_matrix_vector"]
    to, fr = _OneCell(k["map1"], 2), _OneCell(k["map2"], 4)
    y = orc.transpose_apply(to, fr, np.array(k["d"])[None, :, :], np.array(k["x"]))
    np.testing.assert_allclose(y, k["answer"], rtol=0, atol=k["tol"])
    # a plausible mistake, applying D instead of D^T, does not even have the right length
    assert len(k["answer"]) == 2 and not np.allclose(y, np.array(k["d"])[:2, :] @ np.array(k["x"])[:2])


def test_dg_known_answer(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["dg_matrix_vector"]
    to, fr = _OneCell(k["map1"], 4), _OneCell(k["map2"], 2)
    y = orc.dg_apply(to, fr, np.array(k["d"])[None, :, :], np.array(k["x"]))
    np.testing.assert_allclose(y, k["answer"], rtol=0, atol=k["tol"])


def test_dg_inc_known_answer_and_accumulation(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["dg_inc_matrix_vector"]
    to, fr = _OneCell(k["map1"], 4), _OneCell(k["map2"], 2)
    d = np.array(k["d"])
    y = orc.dg_inc_apply(to, fr, d[None, :, :], np.array(k["x"]), np.array(k["y0"]))
    np.testing.assert_allclose(y, k["answer"], rtol=0, atol=k["tol"])
    # GH_READWRITE: adds onto what lhs holds (a zeroing kernel, dg_matrix_vector, would lose y0)
    y0 = np.array([1.0, -2.0, 0.5, 3.0])
    y = orc.dg_inc_apply(to, fr, d[None, :, :], np.array(k["x"]), y0)
    np.testing.assert_allclose(y, y0 + d @ np.array(k["x"]), rtol=0, atol=1e-14)
    assert not np.allclose(orc.dg_apply(to, fr, d[None, :, :], np.array(k["x"])), y)


def test_dg_inc_and_enforce_bc_bitwise_equal_compiled_reference_kernels(orc):
    R = _ref_lib()
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    I = ctypes.c_int
    H = orc.Helmholtz(4, 3)
    w2, w3 = H.w2, H.w3
    x2 = orc.gen_uniform(w2.undf, 1, 3)
    y0 = orc.gen_uniform(w3.undf, 2, 3)
    y3 = orc.dg_inc_apply(w3, w2, H.D, x2, y0)                     # y0 + D x: W2 -> W3
    yr3 = y0.copy()
    R.ref_dg_inc_apply_r64(I(w3.ncells), I(w3.nlayers), I(w3.ndf), I(w3.undf), P(w3.dofmap), I(w2.ndf),
                           I(w2.undf), P(w2.dofmap), P(orc.flat_matrix(H.D)), P(x2), P(yr3))
    np.testing.assert_array_equal(y3, yr3)
    bv = orc.boundary_dofs(orc.W2, 0, 6)
    f = orc.gen_uniform(w2.undf, 3, 3)
    fo = orc.enforce_bc(w2, bv, f)
    fr = f.copy()
    bvf = np.ascontiguousarray(bv.T).ravel()
    R.ref_enforce_bc_r64(I(w2.ncells), I(w2.nlayers), I(w2.ndf), I(w2.undf), P(w2.dofmap), P(bvf), P(fr))
    np.testing.assert_array_equal(fo, fr)


def _w2_surface_dofs(m):
    """0-based dofs of the lowest-order W2 space on the domain's bottom and top surfaces: the
    first and last dof of each vertical-flux column (faces B and T share one column of nl+1)."""
    base = m.dofmap[:, 4] - 1
    assert np.array_equal(m.dofmap[:, 5] - 1, base + 1)
    return np.sort(np.concatenate([base, base + m.nlayers]))


def test_enforce_bc_zeroes_exactly_the_w2_surface_dofs(orc):
    m = orc.Mesh(4, 3, orc.W2, 0)
    bv = orc.boundary_dofs(orc.W2, 0, 6)
    assert bv.shape == (6, 2) and bv[4, 0] == 0 and bv[5, 1] == 0 and bv.sum() == 10
    f = orc.gen_uniform(m.undf, 5, 3) + 2.0  # no zero entries
    g = orc.enforce_bc(m, bv, f)
    surf = _w2_surface_dofs(m)
    assert np.all(g[surf] == 0.0)
    rest = np.setdiff1d(np.arange(m.undf), surf)
    np.testing.assert_array_equal(g[rest], f[rest])
    assert len(surf) == 2 * m.ncells
    # idempotent, and W0 / W3 carry no boundary flags
    np.testing.assert_array_equal(orc.enforce_bc(m, bv, g), g)
    for sp, nd in ((orc.W0, 8), (orc.W3, 1)):
        assert orc.boundary_dofs(sp, 0, nd).min() == 1
    # a planted mistake: the flag on a side face (W) instead of the bottom face fails the check
    bad = bv.copy()
    bad[4, 0], bad[0, 0] = 1, 0
    gb = orc.enforce_bc(m, bad, f)
    assert not np.all(gb[surf] == 0.0)


def test_mm_diagonal_known_answer(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["mm_diagonal"]

    class TwoCells:
        ncells, nlayers, ndf, undf = 2, 1, 3, k["undf"]
        dofmap = np.array(k["map"], np.int32)

    A = np.array([k["mm"], k["mm"]])
    d = orc.mm_diagonal(TwoCells, A)
    np.testing.assert_allclose(d, k["answer"], rtol=0, atol=k["tol"])


# ---- solvers ----------------------------------------------------------------------------------
def test_cg_solves_small_spd_system(orc):
    m = orc.Mesh(4, 3)
    A = orc.gen_spd(8, m.ncell_3d, 0)
    b = orc.gen_uniform(m.undf, 0, 3)
    d = orc.mm_diagonal(m, A)
    x, res, hist = orc.cg(m, A, b, diag=d, r_tol=1e-13, max_iter=m.undf)
    S = orc.assembled_csr(m, A).toarray()
    xs = np.linalg.solve(S, b)
    assert res.converged == 1 and res.iterations < m.undf
    np.testing.assert_allclose(x, xs, rtol=0, atol=1e-10 * np.abs(xs).max())
    xsc, info = spla.cg(S, b, rtol=1e-13, maxiter=1000)
    np.testing.assert_allclose(x, xsc, atol=1e-8 * np.abs(xs).max())


def test_cg_without_monitor_runs_max_iter(orc):
    m = orc.Mesh(4, 3)
    A = orc.gen_spd(8, m.ncell_3d, 0)
    b = orc.gen_uniform(m.undf, 0, 3)
    x, res, _ = orc.cg(m, A, b, diag=orc.mm_diagonal(m, A), max_iter=7, monitor=False)
    assert res.iterations == 7 and res.converged == 0


def test_givens_least_squares_matches_lstsq(orc):
    rng = np.random.default_rng(5)
    mdim = 6
    h = np.triu(rng.standard_normal((mdim + 1, mdim)), -1)
    g = np.zeros(mdim + 1)
    g[0] = 2.5
    u = orc.givens_lsq(h, g)
    np.testing.assert_allclose(u, np.linalg.lstsq(h, g, rcond=None)[0], rtol=1e-12, atol=1e-13)


def test_gmres_solves_nonsymmetric_system(orc):
    rng = np.random.default_rng(2)
    n = 40
    M = np.eye(n) * 4 + rng.standard_normal((n, n)) * 0.3
    b = rng.standard_normal(n)
    x, res, hist = orc.gmres_callbacks(n, lambda v: M @ v, lambda v: v / np.diag(M), b, gcrk=10,
                                       r_tol=1e-12, max_iter=30)
    assert res.converged == 1
    np.testing.assert_allclose(x, np.linalg.solve(M, b), rtol=0, atol=1e-10)
    # with gcrk >= n one restart cycle solves it
    x1, res1, _ = orc.gmres_callbacks(n, lambda v: M @ v, lambda v: v, b, gcrk=n, r_tol=1e-10, max_iter=1)
    assert res1.converged == 1 and res1.iterations == 1


# ---- mixed W2/W3 reduction -------------------------------------------------------------------
def test_helmholtz_composition_matches_assembled_sparse_operators(orc):
    H = orc.Helmholtz(4, 3)
    D = orc.assembled_csr(H.w2, H.D, to_mesh=H.w3)          # W3 x W2
    M3 = orc.assembled_csr(H.w3, H.M3)
    import scipy.sparse as sp
    Hs = (M3 + H.tau * D @ sp.diags(1.0 / H.m2d) @ D.T).toarray()
    Hd = H.dense()
    np.testing.assert_allclose(Hd, Hs, rtol=1e-13, atol=1e-13 * np.abs(Hs).max())
    assert np.allclose(Hd, Hd.T, rtol=1e-12, atol=1e-13)
    assert np.linalg.eigvalsh(Hd).min() > 0.9  # M3 >= 1 on the diagonal, D-part PSD
    # W2 divergence signs W,S,E,N,B,T = -,-,+,+,-,+
    assert np.all(np.sign(H.D[:, 0, :]) == orc.D_SIGNS)


# ---- the remaining solvers of the mass-matrix solver factory ----------------------------------
def _small_system(seed=7, n=30, sym=False):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n)) * 0.3
    M = M @ M.T + n * 0.25 * np.eye(n) if sym else np.eye(n) * 4 + M
    d = np.diag(M).copy()
    b = rng.standard_normal(n)
    return M, d, b


def test_bicgstab_solves_nonsymmetric_system(orc):
    M, d, b = _small_system()
    x, res, _ = orc.solve_callbacks("bicgstab", len(b), lambda v: M @ v, lambda v: v / d, b, r_tol=1e-12,
                                    max_iter=200)
    assert res.converged == 1
    np.testing.assert_allclose(x, np.linalg.solve(M, b), rtol=0, atol=1e-9)


def _right_prec_krylov_min(M, d, b, k, x0=None):
    """Minimal-residual x over x0 + span{P^-1 r0, P^-1 (A P^-1) r0, ..., P^-1 (A P^-1)^(k-1) r0},
    r0 = b - A x0 (P = diag d), by least squares on an explicit basis."""
    x0 = np.zeros_like(b) if x0 is None else x0
    r0 = b - M @ x0
    K = [r0 / d]
    for _ in range(k - 1):
        K.append((M @ K[-1]) / d)
    B = np.array(K).T
    c = np.linalg.lstsq(M @ B, r0, rcond=None)[0]
    return x0 + B @ c


@pytest.mark.parametrize("m", [3, 6])
def test_fgmres_and_gcr_one_cycle_are_minimal_residual(orc, m):
    M, d, b = _small_system(n=25)
    P = lambda v: v / d
    xf, rf, _ = orc.solve_callbacks("fgmres", len(b), lambda v: M @ v, P, b, gcrk=m, max_iter=1, r_tol=1e-30)
    np.testing.assert_allclose(xf, _right_prec_krylov_min(M, d, b, m), rtol=0, atol=1e-10)
    # GCR(m) starts from x0 = P^-1 b and minimises over x0 + P^-1 K_m(A P^-1, b - A x0)
    xg, rg, _ = orc.solve_callbacks("gcr", len(b), lambda v: M @ v, P, b, gcrk=m, max_iter=1, r_tol=1e-30)
    np.testing.assert_allclose(xg, _right_prec_krylov_min(M, d, b, m, x0=b / d), rtol=0, atol=1e-10)
    # restarted to convergence both reach the solution
    for meth in ("fgmres", "gcr"):
        x, r, _ = orc.solve_callbacks(meth, len(b), lambda v: M @ v, P, b, gcrk=m, max_iter=200, r_tol=1e-12)
        assert r.converged == 1
        np.testing.assert_allclose(x, np.linalg.solve(M, b), rtol=0, atol=1e-9)


def test_jacobi_matches_closed_form(orc):
    M, d, b = _small_system(sym=True)
    xs = np.linalg.solve(M, b)
    x0 = np.random.default_rng(1).standard_normal(len(b))
    rho, k = 0.7, 12
    x, res, _ = orc.solve_callbacks("jacobi", len(b), lambda v: M @ v, lambda v: v / d, b, x0=x0, max_iter=k,
                                    monitor=False, rho_relax=rho)
    G = np.eye(len(b)) - rho * (M / d[:, None])
    np.testing.assert_allclose(x, xs + np.linalg.matrix_power(G, k) @ (x0 - xs), rtol=0, atol=1e-12)
    x, res, _ = orc.solve_callbacks("jacobi", len(b), lambda v: M @ v, lambda v: v / d, b, x0=x0, max_iter=2000,
                                    r_tol=1e-10, rho_relax=rho)
    assert res.converged == 1 and res.r_nrm <= 1e-10 * res.r_nrm_0


def test_chebyshev_converges_with_exact_bounds_and_is_linear(orc):
    M, d, b = _small_system(sym=True)
    Dh = 1 / np.sqrt(d)
    ev = np.linalg.eigvalsh(Dh[:, None] * M * Dh[None, :])
    kw = dict(lmin=ev.min(), lmax=ev.max(), max_iter=80)
    A, P = (lambda v: M @ v), (lambda v: v / d)
    x, res, _ = orc.solve_callbacks("chebyshev", len(b), A, P, b, **kw)
    np.testing.assert_allclose(x, np.linalg.solve(M, b), rtol=0, atol=1e-8)
    b2 = np.random.default_rng(9).standard_normal(len(b))
    x2, _, _ = orc.solve_callbacks("chebyshev", len(b), A, P, b2, **kw)
    x12, _, _ = orc.solve_callbacks("chebyshev", len(b), A, P, b + 2 * b2, **kw)
    np.testing.assert_allclose(x12, x + 2 * x2, rtol=0, atol=1e-12)


def test_prec_only_is_one_division(orc):
    M, d, b = _small_system()
    x, res, _ = orc.solve_callbacks("prec_only", len(b), lambda v: M @ v, lambda v: v / d, b)
    np.testing.assert_array_equal(x, b / d)
    assert abs(res.r_nrm - np.linalg.norm(b - M @ x)) <= 1e-13 * np.linalg.norm(b)


# ---- per-cell dense inverse (invert_local_operator_code / matrix_invert_lu) -------------------
def test_invert_local_known_answers(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())
    t = k["invert_local_operator"]
    A = np.full((6 * t["nlayers"], t["ndf"], t["ndf"]), t["value"])
    inv = orc.invert_local(A)
    assert abs(inv[0, 0, 0] - t["answer"]) <= t["tol"]
    mi = k["matrix_invert"]
    a = np.array(mi["a"])
    ai = orc.invert_local(a[None])[0]
    np.testing.assert_allclose(a @ ai, np.eye(4), rtol=0, atol=mi["tol"])


@pytest.mark.parametrize("n", [1, 2, 6, 8, 27])
def test_invert_local_matches_numpy_inverse(orc, n):
    """Nonsymmetric, diagonally dominant elements (LU without pivoting is stable on them): the
    restatement agrees with numpy.linalg.inv, so a transposed or mis-indexed result fails."""
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (5, n, n)) + n * np.eye(n)
    inv = orc.invert_local(A)
    for e in range(5):
        np.testing.assert_allclose(inv[e], np.linalg.inv(A[e]), rtol=1e-12, atol=1e-13)
    inv32 = orc.invert_local(A, r32=True)
    np.testing.assert_allclose(inv32, inv, rtol=2e-5, atol=2e-6)


def test_invert_local_bitwise_equals_compiled_reference(orc):
    R = _ref_lib()
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    for n in (1, 6, 8, 27):
        A = np.random.default_rng(10 + n).uniform(-1, 1, (7, n, n)) + n * np.eye(n)
        for r32, fn, dt in ((False, R.ref_invert_local_r64, np.float64), (True, R.ref_invert_local_r32, np.float32)):
            Af = np.ascontiguousarray(A.astype(dt).transpose(2, 1, 0)).ravel()
            out = np.zeros_like(Af)
            fn(ctypes.c_int64(7), ctypes.c_int(n), P(Af), P(out))
            ref = out.reshape((n, n, 7)).transpose(2, 1, 0)
            np.testing.assert_array_equal(orc.invert_local(A, r32=r32), ref)


# ---- lowest-order W0 mass matrix from coordinates (compute_mass_matrix_w_scalar_code) ---------
_VX = np.array([0, 1, 1, 0, 0, 1, 1, 0])
_VY = np.array([0, 0, 1, 1, 0, 0, 1, 1])
_VZ = np.array([0, 0, 0, 0, 1, 1, 1, 1])


def _m_ref():
    m1 = np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])
    return m1[_VX][:, _VX] * m1[_VY][:, _VY] * m1[_VZ][:, _VZ]


def _one_cell(X):
    """One cell, one layer: vertex columns 1..4 with bottom/top dofs (1,2), (3,4), ..."""
    dm = np.array([[1, 3, 5, 7, 2, 4, 6, 8]], np.int32)
    chi = np.zeros((3, 8))
    for v in range(8):
        chi[:, dm[0, v] - 1] = X[v]
    return dm, chi


def test_gauss_rule_matches_numpy_and_the_reference_weights(orc):
    for n in range(1, 9):
        p, w = orc.gauss_rule(n)
        x, wl = np.polynomial.legendre.leggauss(n)
        np.testing.assert_allclose(p, 0.5 * (x + 1), rtol=0, atol=4e-16)
        np.testing.assert_allclose(w, 0.5 * wl, rtol=0, atol=2e-15)  # weights from the last Newton step
    # get_unit_test_q3x3x3_quadrature_mod.f90: 3x3 horizontal weights, 15 digits
    p, w = orc.gauss_rule(3)
    np.testing.assert_allclose(np.outer(w, w).ravel()[:3], [0.077160493827160, 0.123456790123456, 0.077160493827160],
                               atol=1e-14)


def test_mass_matrix_w0_known_answer(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["mass_matrix_w0"]
    nl, dx, dy, dz = k["nlayers"], k["dx"], k["dy"], k["dz"]
    verts = np.array(k["vertices_of_cell"])
    dm = np.zeros((9, 8), np.int32)
    dm[:, :4] = verts * (nl + 1) + 1
    dm[:, 4:] = verts * (nl + 1) + 2
    undf = 9 * (nl + 1)
    chi = np.zeros((3, undf))
    cell = 0
    for j in range(3):
        for i in range(3):
            for kk in range(nl + 1):
                chi[:, dm[cell, 0] - 1 + kk] = (i * dx, j * dy, kk * dz)
            cell += 1
    mm = orc.mass_matrix_w0(9, nl, dm, chi[0], chi[1], chi[2], nqp=3)
    ans = np.array(k["answer_column_major"]).reshape((8, 8), order="F")
    vol = dx * dy * dz
    np.testing.assert_allclose(mm[0] / vol, ans / vol, rtol=0, atol=k["tol"])
    np.testing.assert_allclose(mm[0], vol * _m_ref(), rtol=1e-14)


def test_mass_matrix_w0_closed_forms(orc):
    rng = np.random.default_rng(3)
    # sheared parallelepiped: constant Jacobian [a b c], M = det * M_ref
    a, b, c = rng.uniform(-1, 1, (3, 3)) + np.diag([3.0, 2.0, 4.0])
    O = rng.uniform(-5, 5, 3)
    X = [O + _VX[v] * a + _VY[v] * b + _VZ[v] * c for v in range(8)]
    dm, chi = _one_cell(X)
    det = np.linalg.det(np.column_stack([a, b, c]))
    for nqp in (2, 3, 5):
        mm = orc.mass_matrix_w0(1, 1, dm, chi[0], chi[1], chi[2], nqp=nqp)[0]
        np.testing.assert_allclose(mm, det * _m_ref(), rtol=1e-13)
    # a general (non-affine) hexahedron: symmetric, positive definite, entries add up to the
    # volume, and 3 Gauss points per direction are already exact
    Xg = [x + rng.uniform(-0.3, 0.3, 3) for x in X]
    dm, chi = _one_cell(Xg)
    m3 = orc.mass_matrix_w0(1, 1, dm, chi[0], chi[1], chi[2], nqp=3)[0]
    m6 = orc.mass_matrix_w0(1, 1, dm, chi[0], chi[1], chi[2], nqp=6)[0]
    np.testing.assert_array_equal(m3, m3.T)
    assert np.all(np.linalg.eigvalsh(m3) > 0)
    np.testing.assert_allclose(m3, m6, rtol=1e-12)
    # the volume from a fine midpoint rule on the trilinear map's Jacobian (independent of the
    # Gauss rule and of the basis ordering)
    n = 40
    t = (np.arange(n) + 0.5) / n
    xs, ys, zs = np.meshgrid(t, t, t, indexing="ij")
    P = np.array(Xg)
    J = np.zeros(xs.shape + (3, 3))
    for v in range(8):
        lx = np.where(_VX[v], xs, 1 - xs); ly = np.where(_VY[v], ys, 1 - ys); lz = np.where(_VZ[v], zs, 1 - zs)
        sx, sy, sz = (1 if _VX[v] else -1), (1 if _VY[v] else -1), (1 if _VZ[v] else -1)
        for r in range(3):
            J[..., r, 0] += P[v, r] * sx * ly * lz
            J[..., r, 1] += P[v, r] * lx * sy * lz
            J[..., r, 2] += P[v, r] * lx * ly * sz
    vol = np.linalg.det(J).mean()
    assert abs(m3.sum() - vol) <= 1e-3 * abs(vol)


def _kat_box_mesh(k):
    """The bi-periodic 3x3 mesh of the mass-matrix tests with the coordinates of
    get_w0_3x3x3_field: each cell's south-west vertex column at (i dx, j dy, k dz).  Cell 1 (the
    one the reference tests check) is then the box dx x dy x dz in every layer; cells on the
    periodic seam are not boxes."""
    nl, dx, dy, dz = k["nlayers"], k["dx"], k["dy"], k["dz"]
    verts = np.array(json.loads((GOLDEN / "kat.json").read_text())["mass_matrix_w0"]["vertices_of_cell"])
    dm = np.zeros((9, 8), np.int32)
    dm[:, :4] = verts * (nl + 1) + 1
    dm[:, 4:] = verts * (nl + 1) + 2
    chi = np.zeros((3, 9 * (nl + 1)))
    for c in range(9):
        for kk in range(nl + 1):
            chi[:, dm[c, 0] - 1 + kk] = ((c % 3) * dx, (c // 3) * dy, kk * dz)
    return dm, chi


# lowest-order W2 dofs W, S, E, N, B, T: component, sign of the reference normal, and whether
# the linear factor is t (1) or 1 - t (0) -- the closed forms below integrate these
_W2_COMP = np.array([0, 1, 0, 1, 2, 2])
_W2_SIGN = np.array([1.0, -1.0, 1.0, -1.0, 1.0, 1.0])
_W2_FAR = np.array([0, 0, 1, 1, 0, 1])


def _w2_reference_integrals():
    """I_ij = integral over the unit cube of v_i v_j (the scalar factors of the basis)."""
    I = np.zeros((6, 6))
    for i in range(6):
        for j in range(6):
            if _W2_COMP[i] == _W2_COMP[j]:
                one_d = 1 / 3 if _W2_FAR[i] == _W2_FAR[j] else 1 / 6
            else:
                one_d = 1 / 4
            I[i, j] = _W2_SIGN[i] * _W2_SIGN[j] * one_d
    return I


def test_w2_basis_matches_the_reference_tables(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["w2_basis_q3x3x3"]
    b, d = orc.w2_basis(3)
    np.testing.assert_allclose(b, np.array(k["basis_column_major"]).reshape(k["shape"], order="F"), rtol=0,
                               atol=1e-14)
    np.testing.assert_array_equal(d, np.array(k["div_column_major"]).reshape(k["div_shape"], order="F"))


def test_mass_matrix_w2_known_answer_and_closed_forms(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["mass_matrix_w2"]
    dm, chi = _kat_box_mesh(k)
    mm = orc.mass_matrix_w2(9, k["nlayers"], dm, chi[0], chi[1], chi[2], nqp=3)
    ans = np.array(k["answer_column_major"]).reshape((6, 6), order="F")
    norm = (k["dx"] + k["dy"] + k["dz"]) / 3
    for layer in range(k["nlayers"]):  # cell 1 in every layer
        np.testing.assert_allclose(mm[layer] * norm, ans * norm, rtol=0, atol=k["tol"])
    # sheared parallelepiped x = A xi + o: (J b_i).(J b_j) / det = G(c_i, c_j) v_i v_j / det with
    # G = A^T A; a transposed Jacobian (A A^T) or a dropped 1/det fails it
    rng = np.random.default_rng(4)
    A = rng.uniform(-1, 1, (3, 3)) + np.diag([3.0, 2.0, 4.0])
    O = rng.uniform(-5, 5, 3)
    X = [O + A @ np.array([_VX[v], _VY[v], _VZ[v]], float) for v in range(8)]
    dmc, chic = _one_cell(X)
    G = A.T @ A
    want = G[_W2_COMP][:, _W2_COMP] * _w2_reference_integrals() / np.linalg.det(A)
    for nqp in (2, 3, 5):
        got = orc.mass_matrix_w2(1, 1, dmc, chic[0], chic[1], chic[2], nqp=nqp)[0]
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14 * np.abs(want).max())
    # a general hexahedron: symmetric bit for bit, positive definite
    Xg = [x + rng.uniform(-0.3, 0.3, 3) for x in X]
    dmc, chic = _one_cell(Xg)
    m = orc.mass_matrix_w2(1, 1, dmc, chic[0], chic[1], chic[2], nqp=3)[0]
    np.testing.assert_array_equal(m, m.T)
    assert np.all(np.linalg.eigvalsh(m) > 0)


def test_mass_matrix_w3_known_answer_and_invariants(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["mass_matrix_w3"]
    dm, chi = _kat_box_mesh(k)
    vol = k["dx"] * k["dy"] * k["dz"]
    mm = orc.mass_matrix_w3(9, k["nlayers"], dm, chi[0], chi[1], chi[2], nqp=3)
    nl = k["nlayers"]
    np.testing.assert_allclose(mm[:nl, 0, 0] / vol, k["answer_over_volume"], rtol=0, atol=k["tol"])
    inv = orc.mass_matrix_w3(9, nl, dm, chi[0], chi[1], chi[2], nqp=3, rehabilitate=False)
    np.testing.assert_allclose(inv[:nl, 0, 0] * vol, 1.0, rtol=1e-14)
    # a general hexahedron: the W3 entry is the cell volume, which is also the sum of the cell's
    # W0 mass entries (pinned independently); dj is quadratic per direction, so 2 Gauss points
    # per direction already integrate it exactly
    rng = np.random.default_rng(6)
    X = [np.array([_VX[v] * 2.0, _VY[v] * 1.5, _VZ[v] * 3.0]) + rng.uniform(-0.3, 0.3, 3) for v in range(8)]
    dmc, chic = _one_cell(X)
    m0 = orc.mass_matrix_w0(1, 1, dmc, chic[0], chic[1], chic[2], nqp=3)[0]
    for nqp in (2, 3, 5):
        m3 = orc.mass_matrix_w3(1, 1, dmc, chic[0], chic[1], chic[2], nqp=nqp)[0, 0, 0]
        assert abs(m3 - m0.sum()) <= 1e-13 * abs(m3)


def test_div_operator_known_answer_and_invariants(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["div_operator"]
    dm, chi = _kat_box_mesh(k)
    d = orc.div_operator(9, k["nlayers"], dm, chi[0], chi[1], chi[2], nqp=3)
    for layer in range(k["nlayers"]):
        np.testing.assert_allclose(d[layer, 0], k["answer"], rtol=0, atol=k["tol"])
    # without rehabilitation each entry is the divergence sign times the W3 entry without
    # rehabilitation (the same sums of w / dj), bit for bit, on any cell
    rng = np.random.default_rng(8)
    X = [np.array([_VX[v], _VY[v], _VZ[v]], float) + rng.uniform(-0.2, 0.2, 3) for v in range(8)]
    dmc, chic = _one_cell(X)
    dn = orc.div_operator(1, 1, dmc, chic[0], chic[1], chic[2], nqp=3, rehabilitate=False)[0, 0]
    m3n = orc.mass_matrix_w3(1, 1, dmc, chic[0], chic[1], chic[2], nqp=3, rehabilitate=False)[0, 0, 0]
    np.testing.assert_array_equal(dn, np.array(k["answer"]) * m3n)


def test_mass_matrix_wtheta_known_answer_and_invariants(orc):
    """compute_mass_matrix_kernel_wtheta_mod_test.pf: entries (1,1) and (1,2) of the first two
    layers of the 6000 x 1000 x 2000 box cell are 1/3 and 1/6 of its volume."""
    k = json.loads((GOLDEN / "kat.json").read_text())["mass_matrix_w3"]  # the same box and mesh
    dm, chi = _kat_box_mesh(k)
    vol = k["dx"] * k["dy"] * k["dz"]
    mm = orc.mass_matrix_wtheta(9, k["nlayers"], dm, chi[0], chi[1], chi[2], nqp=3)
    for ik in (0, 1):
        np.testing.assert_allclose(mm[ik, 0, 0] / vol, 1 / 3, rtol=0, atol=1e-6)
        np.testing.assert_allclose(mm[ik, 0, 1] / vol, 1 / 6, rtol=0, atol=1e-6)
    # sheared parallelepiped: det(A) [[1/3, 1/6], [1/6, 1/3]]; a general hexahedron: symmetric,
    # entries adding up to the W3 entry (the volume)
    rng = np.random.default_rng(9)
    A = rng.uniform(-1, 1, (3, 3)) + np.diag([3.0, 2.0, 4.0])
    X = [A @ np.array([_VX[v], _VY[v], _VZ[v]], float) for v in range(8)]
    dmc, chic = _one_cell(X)
    got = orc.mass_matrix_wtheta(1, 1, dmc, *chic, nqp=3)[0]
    np.testing.assert_allclose(got, np.linalg.det(A) * np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]]), rtol=1e-12)
    Xg = [x + rng.uniform(-0.3, 0.3, 3) for x in X]
    dmc, chic = _one_cell(Xg)
    m = orc.mass_matrix_wtheta(1, 1, dmc, *chic, nqp=3)[0]
    m3 = orc.mass_matrix_w3(1, 1, dmc, *chic, nqp=3)[0, 0, 0]
    np.testing.assert_array_equal(m, m.T)
    assert abs(m.sum() - m3) <= 1e-13 * abs(m3)


def test_r_solver_cg_converges_like_fp64_cg(orc):
    """The real32 r_solver CG restatement against scipy's FP64 solve of the same (real32-rounded)
    system and the FP64 oracle CG: same solution to real32 accuracy, iteration count within one
    of the FP64 count at a tolerance real32 can reach.  A dropped real32 rounding of alpha or a
    wrong sign in the residual update fails the solution check."""
    m = orc.Mesh(4, 3)
    A32 = orc.gen_spd(8, m.ncell_3d, 0).astype(np.float32).astype(np.float64)
    b = orc.gen_uniform(m.undf, 0, 3).astype(np.float32).astype(np.float64)
    diag = orc.mm_diagonal(m, A32)
    ref = spla.spsolve(orc.assembled_csr(m, A32).tocsc(), b)
    x32, it32, ok = orc.cg_r32(m, A32, b, diag.astype(np.float32), r_tol=1e-5)
    assert ok
    assert np.linalg.norm(x32 - ref) <= 1e-4 * np.linalg.norm(ref)
    _, res64, _ = orc.cg(m, A32, b, diag=diag, r_tol=1e-5, a_tol=0.0)
    assert abs(it32 - res64.iterations) <= 1


def test_extrusion_eta_pinned(orc):
    k = json.loads((GOLDEN / "kat.json").read_text())["extrusion"]
    for kind, name in ((0, "uniform"), (1, "quadratic"), (2, "geometric")):
        np.testing.assert_allclose(orc.eta(kind, k["layers"]), k[name], rtol=0, atol=k["tol"])
    for nl in (1, 5, 70):
        kk = np.arange(nl + 1) / nl
        assert np.array_equal(orc.eta(0, nl), kk)
        assert np.array_equal(orc.eta(1, nl), kk * kk)
        g = orc.eta(2, nl)
        assert g[0] == 0.0 and abs(g[-1] - 1.0) < 1e-12 and np.all(np.diff(g) > 0)
        # geometric layers grow by the stretching factor 1.03
        np.testing.assert_allclose(np.diff(g)[1:] / np.diff(g)[:-1], 1.03, rtol=1e-9)


def test_cubedsphere_node_coordinates_pinned_to_mesh_files(orc):
    """The mesh generator's vertex longitudes and latitudes against the reference's own unrotated
    unit-test meshes (degrees as the files hold them), and on the unit sphere at the equiangular
    grid angles."""
    ref = np.load(GOLDEN / "node_coords.npz")
    for n in (4, 8, 16, 32):
        lon, lat = orc.cubedsphere_lonlat(n)
        dx = (np.degrees(lon) - ref[f"C{n}_node_x"] + 180.0) % 360.0 - 180.0
        assert np.abs(dx).max() <= 1e-10, n
        assert np.abs(np.degrees(lat) - ref[f"C{n}_node_y"]).max() <= 1e-10, n
    # panel I's vertices sit at (1, tan a, tan b) / |.| on the grid a, b = -pi/4 + j pi/(2n)
    n = 6
    lon, lat = orc.cubedsphere_lonlat(n)
    xyz = np.stack([np.cos(lon) * np.cos(-lat), np.sin(lon) * np.cos(-lat), np.sin(-lat)], 1)
    a = -np.pi / 4 + np.arange(n) * np.pi / (2 * n)
    t1, t2 = np.meshgrid(np.tan(a), np.tan(a))
    p1 = np.stack([np.ones_like(t1), t1, t2], -1).reshape(-1, 3)
    p1 /= np.linalg.norm(p1, axis=1)[:, None]
    np.testing.assert_allclose(xyz[: n * n], p1, atol=1e-14)
    # a planted mistake, latitude not negated, moves every off-equator vertex
    assert np.abs(np.sin(lat[: n * n]) + p1[:, 2]).max() < 1e-14


def test_coordinate_fields_on_shells(orc):
    m = orc.Mesh(5, 6)
    top, R = 80000.0, 6371229.0
    for kind in (0, 1, 2):
        chi = orc.chi_w0(m, kind, top, R)
        r = np.sqrt(chi[0] ** 2 + chi[1] ** 2 + chi[2] ** 2)
        e = orc.eta(kind, m.nlayers)
        lon, lat = orc.cubedsphere_lonlat(5)
        for c in (0, 17, 149):
            for i in range(4):
                base = m.dofmap[c, i] - 1
                np.testing.assert_allclose(r[base:base + m.nlayers + 1], R + e * top, rtol=1e-14)
                g = m.vert_gid[c, i] - 1
                u = np.array([np.cos(lon[g]) * np.cos(lat[g]), np.sin(lon[g]) * np.cos(lat[g]), np.sin(lat[g])])
                np.testing.assert_allclose(np.stack(chi, 1)[base:base + m.nlayers + 1] / r[base:base + m.nlayers + 1, None],
                                           np.tile(u, (m.nlayers + 1, 1)), atol=1e-14)
