"""Function spaces built from UGRID mesh files (the reference's own files, tests/golden/ugrid/).

Bars: on the C12 cubed-sphere file, the dofmap and colouring of every space equal those of the
generated C12 mesh (the file is that mesh) bit for bit; on the bi-periodic 8x8 planar file, the
dofmap equals the oracle's (local vertex ids, connectivity and dofmap_setup restated from the
file's connectivity) and the colouring equals the oracle's planar breadth-first colouring; on
the bounded planar file (no npanels attribute) the colouring is the oracle's generic one.  An
LMA apply on each planar space agrees with the oracle to the rounding of its terms."""
import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
UGRID = GOLDEN / "ugrid"


@pytest.mark.parametrize("space,order", [("W0", 0), ("W2", 0), ("W3", 0), ("W0", 1)])
def test_ugrid_cubedsphere_equals_generated(hip, space, order):
    sp = getattr(hip, space)
    a = hip.FunctionSpace.from_ugrid(UGRID / "mesh_C12.nc", 5, sp, order)
    b = hip.FunctionSpace.cubedsphere(12, 5, sp, order)
    assert (a.ncells, a.ndf, a.undf, a.ncolours) == (b.ncells, b.ndf, b.undf, b.ncolours)
    np.testing.assert_array_equal(a.dofmap(), b.dofmap())
    for x, y in zip(a.colouring(), b.colouring()):
        np.testing.assert_array_equal(x, y)


def _oracle_space(orc, fn, fl, nl, space, order=0):
    lid, nv = orc.local_vert_lids(fn)
    foc, eoc, nf, ne = orc.mesh_connectivity(fl, lid)
    return orc.dofmap(space, order, nl, lid, foc, eoc, nv, ne)


@pytest.mark.parametrize("fname,colouring", [("mesh_BiP8x8-750x250.nc", "planar"), ("mesh_planar.nc", "generic")])
def test_ugrid_planar_spaces_match_oracle(hip, orc, fname, colouring):
    from lfric_core_amd.fem import read_ugrid
    fn, fl, _ = read_ugrid(UGRID / fname)
    nl = 4
    for sp, osp in ((hip.W0, orc.W0), (hip.W2, orc.W2), (hip.W3, orc.W3)):
        fs = hip.FunctionSpace.from_ugrid(UGRID / fname, nl, sp)
        dm, undf = _oracle_space(orc, fn, fl, nl, osp)
        assert fs.undf == undf
        np.testing.assert_array_equal(fs.dofmap(), dm)
    cm, npc, cic = orc.colour_planar(fl) if colouring == "planar" else orc.colour_generic(fl)
    got_npc, got_cic = fs.colouring()
    np.testing.assert_array_equal(got_npc, npc)
    np.testing.assert_array_equal(got_cic, cic)
    # an LMA apply on the planar W0 space equals the oracle's on the same mesh
    fs0 = hip.FunctionSpace.from_ugrid(UGRID / fname, nl, hip.W0)
    dm0, undf0 = _oracle_space(orc, fn, fl, nl, orc.W0)
    A = orc.gen_spd(8, fs0.ncell_3d, 3)
    op = hip.LmaOperator(fs0).upload(A)
    x = np.random.default_rng(4).uniform(-1, 1, undf0)
    y = hip.FieldVector(fs0)
    op.apply(hip.FieldVector(fs0).import_field(x), y)

    class _M:  # the oracle's view of this mesh for orc.lma_apply
        pass

    m = _M()
    m.ncells, m.nlayers, m.ndf, m.undf, m.dofmap = fn.shape[0], nl, 8, undf0, dm0
    m.ncells_per_colour, m.cells_in_colour = npc, cic
    # the colour-ordered apply contracts into FMAs, so it agrees with the oracle to rounding of
    # the terms (the bound of the apply parity tests), not bit for bit
    bound = 1e-13 * (abs(orc.assembled_csr(m, A)) @ np.abs(x))
    assert np.all(np.abs(y.export_field() - orc.lma_apply(m, A, x)) <= bound)
