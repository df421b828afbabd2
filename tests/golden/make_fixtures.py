"""Write the golden fixtures under tests/golden/ from data files and test tables held by the
lfric_core reference tree.  Run once in a container that has the reference mounted:

    python tests/golden/make_fixtures.py /root/reference

Everything written is data (inputs and expected outputs); no reference source text is kept.

* mesh_fixtures.npz   face_node / face_link connectivity of the committed UGRID meshes
                      (infrastructure/unit-test/data/mesh_C{4,8,16,32}.nc,
                      applications/coupled/example/mesh_C12.nc,
                      applications/io_demo/example/mesh_C24.nc), read with scipy's netCDF3
                      reader; the C16/C24/C32 arrays are stored as SHA-256 digests only.
* node_coords.npz     node longitudes / latitudes (degrees) of the unrotated unit-test cubed spheres
                      infrastructure/unit-test/data/mesh_C{4,8,16,32}.nc (north pole at 90 N).
* ugrid/*.nc          copies of four of the reference's UGRID mesh files (C4, C12 cubed spheres,
                      a bi-periodic 8x8 planar mesh, a bounded planar mesh): data the reader parses.
* canned_mesh.json    the bi-periodic 3x3x3 unit-test mesh of
                      infrastructure/source/mesh/mesh_mod.F90:3625-3777 (layer-1 cell_next,
                      vert_on_cell, edge_on_cell, face_on_cell) and the W0/W1/W2/W3 dofmaps
                      that infrastructure/unit-test/support/fs_test_data_mod.F90 expects on
                      it (fs0k0_dofmap :24-33, fs1k0_dofmap :346-359, fs2k0_dofmap :1026-1033,
                      fs3k0_dofmap :1824-1826).
* kat.json            the matrix_vector known-answer test
                      (components/science/unit-test/kernel/algebra/matrix_vector_mod_test.pf:57-89),
                      the transpose / dg / dg_inc / mm_diagonal and invert_local_operator kernel
                      tests, the uniform / quadratic / geometric extrusions
                      (infrastructure/unit-test/mesh/extrusion_mod_test.pf),
                      the matrix_invert_lu test
                      (infrastructure/unit-test/utilities/matrix_invert_mod_test.pf:25,39-46) and
                      the W0, W2 and W3 mass-matrix and weak-divergence tests
                      (components/science/unit-test/kernel/fem/compute_mass_matrix_kernel_w{0,2,3}_mod_test.pf,
                      compute_div_operator_kernel_mod_test.pf), the tridiagonal-solve test
                      (components/science/unit-test/kernel/solver/sci_tri_solve_kernel_mod_test.pf),
                      the W2 basis tables at the
                      3x3x3 Gauss points (components/science/unit-test/support/get_unit_test_q3x3x3_basis_mod.f90)
                      and the C3 mesh-generator summary counts
                      (rose-stem/app/check_summary/file/summary_cubedsphere_c3_ugrid.kgo.txt:8-10).
"""
import hashlib
import json
import re
import sys
from pathlib import Path

import numpy as np
from scipy.io import netcdf_file

HERE = Path(__file__).resolve().parent

MESHES = {
    4: "infrastructure/unit-test/data/mesh_C4.nc",
    8: "infrastructure/unit-test/data/mesh_C8.nc",
    12: "applications/coupled/example/mesh_C12.nc",
    16: "infrastructure/unit-test/data/mesh_C16.nc",
    24: "applications/io_demo/example/mesh_C24.nc",
    32: "infrastructure/unit-test/data/mesh_C32.nc",
}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def read_mesh(path):
    nc = netcdf_file(path, "r", mmap=False)
    pre = [k for k in nc.variables if k.endswith("_face_nodes")][0][: -len("_face_nodes")]
    fn = np.array(nc.variables[pre + "_face_nodes"][:], dtype=np.int32)
    fl = np.array(nc.variables[pre + "_face_links"][:], dtype=np.int32)
    nc.close()
    return fn, fl


def fortran_int_table(text, name):
    """Integers of `integer ... :: name(r,0:c) = reshape( [ ... ], [r,c] )`."""
    m = re.search(re.escape(name) + r"\s*\(([^)]*)\)\s*=\s*reshape\(\s*\[(.*?)\]", text, re.S)
    vals = [int(v) for v in re.findall(r"-?\d+", m.group(2))]
    return vals


def canned_rows(text, start_marker, array, ncells=9):
    i = text.index("self%" + start_marker)
    rows = {}
    for m in re.finditer(r"self%" + array + r"\(:\s*,\s*(\d+)\)\s*=\s*\[([^\]]*)\]", text[i:]):
        c = int(m.group(1))
        if c > ncells or c in rows:
            if len(rows) == ncells:
                break
            continue
        rows[c] = [int(v) for v in re.findall(r"-?\d+", m.group(2))]
    return [rows[c] for c in range(1, ncells + 1)]


UGRID_FILES = (
    "infrastructure/unit-test/data/mesh_C4.nc",
    "applications/coupled/example/mesh_C12.nc",
    "infrastructure/unit-test/data/mesh_BiP8x8-750x250.nc",
    "infrastructure/unit-test/data/mesh_planar.nc",
)


def main(ref):
    ref = Path(ref)
    arrays, digests = {}, {}
    for n, rel in MESHES.items():
        fn, fl = read_mesh(ref / rel)
        digests[f"C{n}_face_nodes"] = digest(fn)
        digests[f"C{n}_face_links"] = digest(fl)
        if n <= 12:
            arrays[f"C{n}_face_nodes"] = fn
            arrays[f"C{n}_face_links"] = fl
    np.savez_compressed(HERE / "mesh_fixtures.npz", **arrays)
    coords = {}
    for n in (4, 8, 16, 32):
        nc = netcdf_file(ref / MESHES[n], "r", mmap=False)
        pre = [k for k in nc.variables if k.endswith("_node_x")][0][: -len("_node_x")]
        coords[f"C{n}_node_x"] = np.array(nc.variables[pre + "_node_x"][:], dtype=np.float64)
        coords[f"C{n}_node_y"] = np.array(nc.variables[pre + "_node_y"][:], dtype=np.float64)
        nc.close()
    np.savez_compressed(HERE / "node_coords.npz", **coords)
    # a few of the reference's own UGRID data files, byte for byte, for the mesh reader's tests
    (HERE / "ugrid").mkdir(exist_ok=True)
    for rel in UGRID_FILES:
        (HERE / "ugrid" / Path(rel).name).write_bytes((ref / rel).read_bytes())
    (HERE / "mesh_digests.json").write_text(json.dumps(digests, indent=1, sort_keys=True) + "\n")

    mesh_src = (ref / "infrastructure/source/mesh/mesh_mod.F90").read_text()
    marker = "else if (mesh_cfg == PLANE_BI_PERIODIC) then"
    # the second occurrence holds the connectivity tables
    first = mesh_src.index(marker)
    start = mesh_src.index(marker, first + 1)
    sub = mesh_src[start:]
    cell_next = [r[:4] for r in canned_rows(sub, "cell_next", "cell_next")]
    vert_on_cell = [r[:4] for r in canned_rows(sub, "vert_on_cell", "vert_on_cell")]
    edge_on_cell = canned_rows(sub, "edge_on_cell", "edge_on_cell")
    face_on_cell = canned_rows(sub, "face_on_cell", "face_on_cell")
    fs = (ref / "infrastructure/unit-test/support/fs_test_data_mod.F90").read_text()
    canned = {
        "ncells": 9,
        "nlayers": 3,
        "cell_next": cell_next,
        "vert_on_cell": vert_on_cell,
        "edge_on_cell": edge_on_cell,
        "face_on_cell": face_on_cell,
        "fs0k0_dofmap": np.array(fortran_int_table(fs, "fs0k0_dofmap")).reshape(10, 8)[1:].tolist(),
        "fs1k0_dofmap": np.array(fortran_int_table(fs, "fs1k0_dofmap")).reshape(10, 12)[1:].tolist(),
        "fs2k0_dofmap": np.array(fortran_int_table(fs, "fs2k0_dofmap")).reshape(10, 6)[1:].tolist(),
        "fs3k0_dofmap": np.array(fortran_int_table(fs, "fs3k0_dofmap")).reshape(10, 1)[1:].tolist(),
    }
    (HERE / "canned_mesh.json").write_text(json.dumps(canned) + "\n")

    # compute_mass_matrix_kernel_w0_mod_test.pf:326-469: cell 1 of the 3x3 mesh of
    # get_unit_test_m3x3_dofmap_mod.f90:37-69, coordinates of get_unit_test_3x3x3_chi_mod.f90:25-62
    # (box dx x dy x dz), answer mass_matrix_w0_r64 (:522-594) normalised by dx*dy*dz
    mm_text = (ref / "components/science/unit-test/kernel/fem/compute_mass_matrix_kernel_w0_mod_test.pf").read_text()
    body = mm_text[mm_text.index("subroutine mass_matrix_w0_r64"):]
    body = body[:body.index("end subroutine mass_matrix_w0_r64")]
    vals = [float(v) for v in re.findall(r"([0-9.]+E[+-][0-9]+)_r_double", body)]
    assert len(vals) == 64
    dims = {k: float(re.search(k + r"\s*=\s*([0-9.]+)_r_def", mm_text).group(1)) for k in ("dx", "dy", "dz")}
    dm_text = (ref / "components/science/unit-test/support/get_unit_test_m3x3_dofmap_mod.f90").read_text()
    dm_body = dm_text[dm_text.index("subroutine get_w0_m3x3_dofmap"):]
    dm_body = dm_body[dm_body.index("reshape"):dm_body.index("[8,9]")]
    table = [int(v) for v in re.findall(r"\b(\d+)\b", dm_body)]
    assert len(table) == 72
    d = np.array([[2.0, -2.0], [3.0, -1.0], [0.4, 1.3], [-4.5, 7.9]])
    x = np.array([0.25, 1.5])
    # compute_mass_matrix_kernel_w2_mod_test.pf:88-183 (answer mass_matrix_w2 :197-), the W3
    # mass test (compute_mass_matrix_kernel_w3_mod_test.pf:215-216, answer 1 after dividing by
    # dx*dy*dz) and compute_div_operator_kernel_mod_test.pf:183-184 (answer div_operator), all on
    # the same 6000 x 1000 x 2000 box cell with the 3x3x3 Gauss rule
    fem = ref / "components/science/unit-test/kernel/fem"
    w2_text = (fem / "compute_mass_matrix_kernel_w2_mod_test.pf").read_text()
    w2_body = w2_text[w2_text.index("subroutine mass_matrix_w2( mm )"):]
    w2_body = w2_body[:w2_body.index("end subroutine mass_matrix_w2")]
    w2_vals = [float(v) for v in re.findall(r"([0-9.]+E[+-][0-9]+)_r_def", w2_body)]
    assert len(w2_vals) == 36
    div_text = (fem / "compute_div_operator_kernel_mod_test.pf").read_text()
    div_body = div_text[div_text.index("subroutine div_operator( div )"):]
    div_body = div_body[div_body.index("d32 = "):div_body.index("shape(d32)")]
    div_body = div_body[div_body.index("["):]
    div_vals = [float(v) for v in re.findall(r"-?\d+", div_body)]
    assert len(div_vals) == 6
    w3_text = (fem / "compute_mass_matrix_kernel_w3_mod_test.pf").read_text()
    w3_answer = float(re.search(r"answer = ([0-9.]+)_r_double", w3_text).group(1))
    # W2 basis and divergence at the 3x3x3 Gauss points (get_unit_test_q3x3x3_basis_mod.f90:1026-1263)
    qb = (ref / "components/science/unit-test/support/get_unit_test_q3x3x3_basis_mod.f90").read_text()

    def basis_table(name, end):
        body = qb[qb.index(name + " = reshape("):]
        body = body[:body.index(end)]
        return [float(v) for v in re.findall(r"(-?[0-9]+\.[0-9]+)_r_def", body)]

    w2_basis = basis_table("basis_w2", "end subroutine get_w2_q3x3x3_basis")
    w2_div = basis_table("diff_basis_w2", "end subroutine get_w2_q3x3x3_diff_basis")
    assert len(w2_basis) == 3 * 6 * 27 and len(w2_div) == 6 * 27
    # sci_tri_solve_kernel_mod_test.pf:29-53: one 10-level W3 column, constant diagonals
    ts = (ref / "components/science/unit-test/kernel/solver/sci_tri_solve_kernel_mod_test.pf").read_text()
    ts_x = [float(v) for v in re.findall(r"-?\d+", re.search(r"x\s*=\s*real\(\(/(.*?)/\)", ts, re.S).group(1))]
    ts_ans = [float(v) for v in re.findall(r"(-?[0-9]+\.[0-9]+)_r_def", re.search(r"answer\s*=(.*?)/\)", ts, re.S).group(1))]
    ts_diag = {n: float(re.search(n + r"\s*=\s*(-?[0-9.]+)_r_def", ts).group(1)) for n in ("A0", "AP", "AM")}
    assert len(ts_x) == 10 and len(ts_ans) == 10
    summary = (ref / "rose-stem/app/check_summary/file/summary_cubedsphere_c3_ugrid.kgo.txt").read_text()
    counts = {k: int(re.search(k + r":\s*(\d+)", summary).group(1)) for k in ("Nodes", "Edges", "Faces")}
    kat = {
        "matrix_vector": {
            "d": d.tolist(), "x": x.tolist(), "answer": [-2.5, -0.75, 2.05, 10.725],
            "map1": [1, 2, 3, 4], "map2": [1, 2], "tol": 1e-14,
        },
        # transpose_matrix_vector_mod_test.pf:33-67: y(2) = D(4x2)^T x(4), answer = matmul
        "transpose_matrix_vector": {
            "d": d.tolist(), "x": [0.25, 1.5, -0.35, 5.0],
            "answer": (d.T @ np.array([0.25, 1.5, -0.35, 5.0])).tolist(),
            "map1": [1, 2], "map2": [1, 2, 3, 4], "tol": 1e-14,
        },
        # dg_matrix_vector_mod_test.pf:38-68: y(4) = D(4x2) x(2), answer = matmul
        "dg_matrix_vector": {
            "d": d.tolist(), "x": x.tolist(), "answer": (d @ x).tolist(),
            "map1": [1, 2, 3, 4], "map2": [1, 2], "tol": 1e-14,
        },
        # extrusion_mod_test.pf:181-210 (uniform), 250-271 (quadratic), 306-328 (geometric): eta(0:5)
        "extrusion": {
            "layers": 5, "tol": 1e-3,
            "uniform": [0.0, 0.2, 0.4, 0.6, 0.8, 1.0],
            "quadratic": [0.0, 0.04, 0.16, 0.36, 0.64, 1.0],
            "geometric": [0.0, 0.188, 0.382, 0.582, 0.788, 1.0],
        },
        # dg_inc_matrix_vector_mod_test.pf:44-85: y(4) = 0 + D(4x2) x(2), answer = matmul
        "dg_inc_matrix_vector": {
            "d": d.tolist(), "x": x.tolist(), "y0": [0.0, 0.0, 0.0, 0.0], "answer": (d @ x).tolist(),
            "map1": [1, 2, 3, 4], "map2": [1, 2], "tol": 1e-14,
        },
        # mm_diagonal_kernel_mod_test.pf:42-84: two cells sharing dof 3, literal answer
        "mm_diagonal": {
            "mm": [[4.0, 1.0, 1.0], [2.0, 6.0, 2.0], [1.0, 1.0, 4.0]],  # mm(cell, i, j), same per cell
            "map": [[1, 2, 3], [3, 4, 5]], "undf": 5,
            "answer": [4.0, 6.0, 8.0, 6.0, 4.0], "tol": 1e-12,
        },
        # invert_local_operator_kernel_mod_test.pf:87-98: W3 (1 dof), 3 layers, every entry 5,
        # inverse entry (1,1,1) = 1/5
        "invert_local_operator": {"ndf": 1, "nlayers": 3, "value": 5.0, "answer": 0.2, "tol": 1e-6},
        # matrix_invert_mod_test.pf:25,39-46: matrix_invert_lu of a 4x4 matrix, a a^-1 = I within tol
        "matrix_invert": {
            "a": (np.array([[4.0, 2.0, 2.0, 1.0], [2.0, 4.0, 1.0, 2.0], [2.0, 1.0, 4.0, 2.0],
                            [1.0, 2.0, 2.0, 4.0]]) / 9.0).tolist(),
            "tol": 1e-7,
        },
        "mass_matrix_w0": {
            "vertices_of_cell": [table[8 * c: 8 * c + 4] for c in range(9)], "nlayers": 3,
            "dx": dims["dx"], "dy": dims["dy"], "dz": dims["dz"],
            "answer_column_major": vals, "tol": 1e-6,
        },
        "mass_matrix_w2": {
            "dx": dims["dx"], "dy": dims["dy"], "dz": dims["dz"], "nlayers": 3,
            "answer_column_major": w2_vals, "normalise": "(dx + dy + dz) / 3", "tol": 1e-6,
        },
        "mass_matrix_w3": {"dx": dims["dx"], "dy": dims["dy"], "dz": dims["dz"], "nlayers": 3,
                           "answer_over_volume": w3_answer, "tol": 1e-6},
        "div_operator": {"dx": dims["dx"], "dy": dims["dy"], "dz": dims["dz"], "nlayers": 3,
                         "answer": div_vals, "tol": 1e-6},
        "w2_basis_q3x3x3": {"basis_column_major": w2_basis, "shape": [3, 6, 9, 3],
                            "div_column_major": w2_div, "div_shape": [1, 6, 9, 3], "digits": 15},
        "tri_solve": {"x": ts_x, "A0": ts_diag["A0"], "AP": ts_diag["AP"], "AM": ts_diag["AM"],
                      "answer": ts_ans, "digits": 15},
        "c3_summary": counts,
    }
    (HERE / "kat.json").write_text(json.dumps(kat, indent=1) + "\n")
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
