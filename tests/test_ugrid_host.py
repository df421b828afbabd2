"""The UGRID mesh reader (host code, no device) and the oracle's planar colouring.

The reader is checked against the reference's own mesh files (tests/golden/ugrid/, copied byte
for byte) and the connectivity arrays extracted from them with scipy's independent netCDF
reader (mesh_fixtures.npz), against files that scipy writes in both classic formats with
0-based indices and fill values, and for clean refusals of what it does not read."""
import json

import numpy as np
import pytest
from scipy.io import netcdf_file

from conftest import GOLDEN

UGRID = GOLDEN / "ugrid"


@pytest.fixture(scope="module")
def reader():
    from lfric_core_amd.fem import read_ugrid
    return read_ugrid


@pytest.mark.parametrize("n", [4, 12])
def test_reader_matches_the_reference_files(reader, orc, n):
    fn, fl, npanels = reader(UGRID / f"mesh_C{n}.nc")
    fx = np.load(GOLDEN / "mesh_fixtures.npz")
    np.testing.assert_array_equal(fn, fx[f"C{n}_face_nodes"])
    np.testing.assert_array_equal(fl, fx[f"C{n}_face_links"])
    assert npanels == 6
    # and the reference's file is the generated cubed sphere
    cn, vc = orc.cubedsphere_topology(n)
    np.testing.assert_array_equal(fn, vc)
    np.testing.assert_array_equal(fl, cn)


def test_reader_on_planar_meshes(reader):
    fn, fl, npanels = reader(UGRID / "mesh_BiP8x8-750x250.nc")
    assert fn.shape == (64, 4) and npanels == 1 and fl.min() >= 1  # bi-periodic: every side linked
    fn, fl, npanels = reader(UGRID / "mesh_planar.nc")
    assert fn.shape == (64, 4) and npanels == 0  # no npanels attribute in this file
    assert (fl == 0).sum() == 4 * 8  # the bounded 8x8 mesh: 32 sides without a neighbour


@pytest.mark.parametrize("version", [1, 2])
def test_reader_on_files_written_by_scipy(reader, tmp_path, version):
    rng = np.random.default_rng(version)
    nf = 7
    nodes = rng.integers(0, 20, (nf, 4)).astype(np.int32)       # 0-based, as start_index says
    links = rng.integers(0, nf, (nf, 4)).astype(np.int32)
    links[2, 1] = -9999                                          # the fill value: no neighbour
    path = tmp_path / f"m{version}.nc"
    f = netcdf_file(str(path), "w", version=version)
    f.createDimension("nmesh_face", nf)
    f.createDimension("Four", 4)
    m = f.createVariable("mesh2d", "i", ())
    m.cf_role = "mesh_topology"
    m.face_node_connectivity = "mesh2d_face_nodes"
    m.face_face_connectivity = "mesh2d_face_links"
    m.npanels = np.int32(1)
    v = f.createVariable("mesh2d_face_nodes", "i", ("nmesh_face", "Four"))
    v[:] = nodes
    v.start_index = np.int32(0)
    v = f.createVariable("mesh2d_face_links", "i", ("nmesh_face", "Four"))
    v[:] = links
    v.start_index = np.int32(0)
    v._FillValue = np.int32(-9999)
    f.close()
    fn, fl, npanels = reader(path)
    np.testing.assert_array_equal(fn, nodes + 1)
    want = links + 1
    want[2, 1] = 0
    np.testing.assert_array_equal(fl, want)
    assert npanels == 1
    fn2, _, _ = reader(path, mesh="mesh2d")
    np.testing.assert_array_equal(fn2, fn)


def test_reader_refusals(reader, tmp_path):
    # the library's error type, as raised (other tests may load the package module afresh)
    LfricError = reader.__globals__["check"].__globals__["LfricError"]
    bad = tmp_path / "not_netcdf.nc"
    bad.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))  # a netCDF-4 (HDF5) signature
    with pytest.raises(LfricError, match="netCDF classic"):
        reader(bad)
    with pytest.raises(LfricError, match="cannot open"):
        reader(tmp_path / "missing.nc")
    with pytest.raises(LfricError, match="no mesh topology"):
        reader(UGRID / "mesh_C4.nc", mesh="no_such_mesh")
    trunc = tmp_path / "truncated.nc"
    trunc.write_bytes((UGRID / "mesh_C4.nc").read_bytes()[:200])
    with pytest.raises(LfricError):
        reader(trunc)


def test_oracle_planar_colouring(reader, orc):
    """set_colours_planar_quadrilateral on the bi-periodic 8x8 mesh: a valid 4-colouring in which
    crossing a W/E side flips the lowest bit of colour-1 and crossing an S/N side the next one;
    refused (None) on a mesh that is not planar quadrilateral."""
    _, fl, _ = reader(UGRID / "mesh_BiP8x8-750x250.nc")
    res = orc.colour_planar(fl)
    assert res is not None
    cm, npc, _ = res
    assert orc.colour_validate(fl, cm) == 1 and cm[0] == 1 and list(npc) == [16, 16, 16, 16]
    for c in range(fl.shape[0]):
        for d, mask in ((0, 1), (2, 1), (1, 2), (3, 2)):  # W, E flip bit 0; S, N flip bit 1
            assert (cm[fl[c, d] - 1] - 1) == ((cm[c] - 1) ^ mask)
    cn, _ = orc.cubedsphere_topology(4)
    assert orc.colour_planar(cn) is None  # panel edges turn the sides round
