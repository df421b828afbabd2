"""ORACLE — test infrastructure only.

ctypes bindings and numpy helpers over liboracle.so, the plain-C CPU restatement of the
LFRic Core LMA-apply and Krylov-solve path (sources and per-function reference citations in
oracle/src/).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker; the product (lfric_core_amd) never does.
"""
import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_c_int_p = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_U64 = ctypes.c_uint64
_I64 = ctypes.c_int64


class OrcProblem(ctypes.Structure):
    _fields_ = [
        ("ncells", _I), ("nlayers", _I), ("ncolours", _I), ("maxcc", _I), ("ndf", _I),
        ("undf", _I), ("map", _VP), ("ncells_per_colour", _VP), ("cells_in_colour", _VP),
        ("matrix", _VP), ("diag", _VP), ("apply_fn", _VP), ("prec_fn", _VP), ("ctx", _VP),
    ]


class OrcResult(ctypes.Structure):
    _fields_ = [("iterations", _I), ("converged", _I), ("r_nrm_0", _D), ("r_nrm", _D)]


CALLBACK = ctypes.CFUNCTYPE(None, _VP, ctypes.POINTER(_D), ctypes.POINTER(_D))

_lib = None


def build():
    """Compile liboracle.so with the committed Makefile (gcc)."""
    import subprocess
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        sig = {
            "orc_cubedsphere_topology": (_I, [_I, _VP, _VP]),
            "orc_local_vert_lids": (_I, [_I, _VP, _VP]),
            "orc_mesh_connectivity": (_I, [_I, _VP, _VP, _VP, _VP, _VP, _VP]),
            "orc_ndf": (_I, [_I, _I]),
            "orc_dofmap": (_I, [_I, _I, _I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP]),
            "orc_colour_cubedsphere": (_I, [_I, _VP, _VP]),
            "orc_colour_generic": (_I, [_I, _VP, _VP]),
            "orc_colour_lookup": (_I, [_I, _I, _VP, _VP, _VP, _I]),
            "orc_colour_validate": (_I, [_I, _VP, _VP]),
            "orc_colour_planar": (_I, [_I, _VP, _VP]),
            "orc_rng_bits": (_U64, [_U64, _U64, _U64]),
            "orc_rng_uniform": (_D, [_U64, _U64, _U64]),
            "orc_rng_normal": (_D, [_U64, _U64, _U64]),
            "orc_log": (_D, [_D]),
            "orc_exp": (_D, [_D]),
            "orc_gen_spd": (None, [_I, _I64, _U64, _I, _VP]),
            "orc_gen_uniform": (None, [_I64, _U64, _U64, _VP]),
            "orc_gen_spd_shift": (None, [_I, _I64, _U64, _I, _D, _VP]),
            "orc_gen_op_uniform": (None, [_I, _I, _I64, _U64, _U64, _D, _D, _VP, _VP]),
            "orc_lma_apply": (None, [_I, _I, _I, _VP, _VP, _I, _I, _I, _VP, _I, _I, _VP, _VP, _VP, _VP]),
            "orc_lma_apply_r32": (None, [_I, _I, _I, _VP, _VP, _I, _I, _I, _VP, _I, _I, _VP, _VP, _VP, _VP]),
            "orc_transpose_apply": (None, [_I, _I, _I, _VP, _VP, _I, _I, _I, _VP, _I, _I, _VP, _VP, _VP, _VP]),
            "orc_dg_apply": (None, [_I, _I, _I, _I, _VP, _I, _I, _VP, _VP, _VP, _VP]),
            "orc_mm_diagonal": (None, [_I, _I, _I, _I, _VP, _VP, _VP]),
            "orc_cubedsphere_lonlat": (_I, [_I, _VP, _VP]),
            "orc_eta": (_I, [_I, _I, _VP]),
            "orc_chi_w0": (_I, [_I, _I, _VP, _VP, _VP, _VP, _I, _D, _D, _VP, _VP, _VP]),
            "orc_dg_inc_apply": (None, [_I, _I, _I, _I, _VP, _I, _I, _VP, _VP, _VP, _VP]),
            "orc_enforce_bc": (None, [_I, _I, _I, _VP, _VP, _VP]),
            "orc_boundary_dofs": (_I, [_I, _I, _I, _VP]),
            "orc_invert_local": (_I, [_I64, _I, _VP, _VP]),
            "orc_invert_local_r32": (_I, [_I64, _I, _VP, _VP]),
            "orc_gauss_rule": (_I, [_I, _VP, _VP]),
            "orc_mass_matrix_w0": (_I, [_I, _I, _VP, _VP, _VP, _VP, _I, _VP]),
            "orc_w2_basis": (_I, [_I, _VP, _VP]),
            "orc_mass_matrix_w2": (_I, [_I, _I, _VP, _VP, _VP, _VP, _I, _VP]),
            "orc_mass_matrix_w3": (_I, [_I, _I, _VP, _VP, _VP, _VP, _I, _I, _VP]),
            "orc_mass_matrix_wtheta": (_I, [_I, _I, _VP, _VP, _VP, _VP, _I, _VP]),
            "orc_div_operator": (_I, [_I, _I, _VP, _VP, _VP, _VP, _I, _I, _VP]),
            "orc_cma_params": (_I, [_I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _VP]),
            "orc_cma_dofmaps": (None, [_I, _I, _I, _I, _VP, _VP, _VP]),
            "orc_tri_solve": (None, [_I, _I, _VP, _VP, _VP, _VP, _VP, _VP]),
            "orc_tri_solve_r32": (None, [_I, _I, _VP, _VP, _VP, _VP, _VP, _VP]),
            "orc_cg": (_I, [ctypes.POINTER(OrcProblem), _VP, _VP, _D, _D, _I, _I, ctypes.POINTER(OrcResult), _VP]),
            "orc_gmres": (_I, [ctypes.POINTER(OrcProblem), _VP, _VP, _I, _D, _D, _I, _I, ctypes.POINTER(OrcResult), _VP]),
            "orc_bicgstab": (_I, [ctypes.POINTER(OrcProblem), _VP, _VP, _D, _D, _I, _I, ctypes.POINTER(OrcResult), _VP]),
            "orc_fgmres": (_I, [ctypes.POINTER(OrcProblem), _VP, _VP, _I, _D, _D, _I, _I, ctypes.POINTER(OrcResult), _VP]),
            "orc_gcr": (_I, [ctypes.POINTER(OrcProblem), _VP, _VP, _I, _D, _D, _I, _I, ctypes.POINTER(OrcResult), _VP]),
            "orc_jacobi": (_I, [ctypes.POINTER(OrcProblem), _VP, _VP, _D, _D, _I, _I, _D, ctypes.POINTER(OrcResult),
                                _VP]),
            "orc_chebyshev": (_I, [ctypes.POINTER(OrcProblem), _VP, _VP, _I, _D, _D, ctypes.POINTER(OrcResult)]),
            "orc_prec_only": (_I, [ctypes.POINTER(OrcProblem), _VP, _VP, ctypes.POINTER(OrcResult)]),
            "orc_givens_lsq": (_I, [_I, _VP, _VP, _VP]),
        }
        for sfx in ("", "_r32"):
            T = ctypes.c_double if sfx == "" else ctypes.c_float
            sig.update({
                "orc_cma_asm" + sfx: (None, [_I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP]),
                "orc_cma_apply" + sfx: (None, [_I, _VP, _VP, _VP, _VP, _I, _VP, _VP, _I, _VP, _VP]),
                "orc_cma_appinv" + sfx: (None, [_I, _VP, _VP, _VP, _VP, _I, _VP, _VP]),
                "orc_cma_mul" + sfx: (None, [_I, _VP, _VP, _VP, _VP, _VP, _VP]),
                "orc_cma_scaledadd" + sfx: (None, [_I, _VP, _VP, _VP, _VP, _VP, _VP, T, T]),
                "orc_cma_asm_diag_hmht" + sfx: (None, [_I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP]),
                "orc_cma_asm_m2v_lumped_inv" + sfx: (None, [_I, _I, _I, _VP, _VP, _VP, _VP]),
            })
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


W0, W1, W2, W3 = 0, 1, 2, 3


def cubedsphere_topology(n):
    nc = 6 * n * n
    cn = np.zeros((nc, 4), np.int32)
    vc = np.zeros((nc, 4), np.int32)
    rc = lib().orc_cubedsphere_topology(n, _p(cn), _p(vc))
    assert rc == 0
    return cn, vc


def local_vert_lids(verts_on_cell):
    lid = np.zeros_like(verts_on_cell, dtype=np.int32)
    nv = lib().orc_local_vert_lids(verts_on_cell.shape[0], _p(np.ascontiguousarray(verts_on_cell, np.int32)), _p(lid))
    return lid, nv


def mesh_connectivity(cell_next, vert_on_cell):
    nc = cell_next.shape[0]
    foc = np.zeros((nc, 6), np.int32)
    eoc = np.zeros((nc, 12), np.int32)
    nf, ne = _I(0), _I(0)
    lib().orc_mesh_connectivity(nc, _p(np.ascontiguousarray(cell_next, np.int32)),
                                _p(np.ascontiguousarray(vert_on_cell, np.int32)), _p(foc), _p(eoc),
                                ctypes.byref(nf), ctypes.byref(ne))
    return foc, eoc, nf.value, ne.value


def dofmap(space, order, nlayers, vert_on_cell, face_on_cell, edge_on_cell, nverts_2d, nedges_2d):
    nc = vert_on_cell.shape[0]
    ndf = lib().orc_ndf(space, order)
    dm = np.zeros((nc, ndf), np.int32)
    undf = _I(0)
    rc = lib().orc_dofmap(space, order, nc, nlayers, nverts_2d, nedges_2d,
                          _p(np.ascontiguousarray(vert_on_cell, np.int32)),
                          _p(np.ascontiguousarray(face_on_cell, np.int32)),
                          _p(np.ascontiguousarray(edge_on_cell, np.int32)), _p(dm), ctypes.byref(undf))
    assert rc == 0, rc
    return dm, undf.value


def colour_cubedsphere(n, cell_next):
    """set_colours for a 6-panel mesh: the closed form, or the generic greedy colouring when the
    closed form does not validate.  Returns (colour_map, ncells_per_colour, cells_in_colour)."""
    nc = 6 * n * n
    cn = _p(np.ascontiguousarray(cell_next, np.int32))
    cm = np.zeros(nc, np.int32)
    ncol = 4
    if lib().orc_colour_cubedsphere(n, cn, _p(cm)) != 0:
        ncol = lib().orc_colour_generic(nc, cn, _p(cm))
        assert ncol > 0, "generic colouring did not validate"
    npc = np.zeros(ncol, np.int32)
    maxcc = int(np.bincount(cm, minlength=ncol + 1)[1:].max())
    cic = np.zeros((ncol, maxcc), np.int32)
    lib().orc_colour_lookup(nc, ncol, _p(cm), _p(npc), _p(cic), maxcc)
    return cm, npc, cic


def _lookup(cm, ncol):
    nc = len(cm)
    npc = np.zeros(ncol, np.int32)
    maxcc = int(np.bincount(cm, minlength=ncol + 1)[1:].max())
    cic = np.zeros((ncol, maxcc), np.int32)
    lib().orc_colour_lookup(nc, ncol, _p(cm), _p(npc), _p(cic), maxcc)
    return npc, cic


def colour_generic(cell_next):
    """set_colours_generic on any mesh: (colour_map, ncells_per_colour, cells_in_colour)."""
    cn = np.ascontiguousarray(cell_next, np.int32)
    cm = np.zeros(cn.shape[0], np.int32)
    ncol = lib().orc_colour_generic(cn.shape[0], _p(cn), _p(cm))
    assert ncol > 0, "generic colouring did not validate"
    return (cm,) + _lookup(cm, ncol)


def colour_planar(cell_next):
    """set_colours_planar_quadrilateral, or None when the reference would fall back to the
    generic colouring."""
    cn = np.ascontiguousarray(cell_next, np.int32)
    cm = np.zeros(cn.shape[0], np.int32)
    if lib().orc_colour_planar(cn.shape[0], _p(cn), _p(cm)) == 0:
        return None
    return (cm,) + _lookup(cm, 4)


def colour_validate(cell_next, colour_map):
    return lib().orc_colour_validate(cell_next.shape[0], _p(np.ascontiguousarray(cell_next, np.int32)),
                                     _p(np.ascontiguousarray(colour_map, np.int32)))


class Mesh:
    """One serial-partition extruded cubed-sphere function space built by the oracle."""

    def __init__(self, n, nlayers, space=W0, order=0):
        self.n, self.nlayers, self.space, self.order = n, nlayers, space, order
        self.cell_next, gvoc = cubedsphere_topology(n)
        self.vert_gid = gvoc
        self.vert_on_cell, self.nverts = local_vert_lids(gvoc)
        self.face_on_cell, self.edge_on_cell, self.nfaces_2d, self.nedges_2d = mesh_connectivity(
            self.cell_next, self.vert_on_cell)
        self.dofmap, self.undf = dofmap(space, order, nlayers, self.vert_on_cell, self.face_on_cell,
                                        self.edge_on_cell, self.nverts, self.nedges_2d)
        self.colour_map, self.ncells_per_colour, self.cells_in_colour = colour_cubedsphere(n, self.cell_next)
        self.ncolours = len(self.ncells_per_colour)
        self.ncells = 6 * n * n
        self.ndf = self.dofmap.shape[1]
        self.ncell_3d = self.ncells * nlayers


def gen_spd(ndf, ncell_3d, seed, scaled=False, shift=0.1):
    """Element matrices in the reference layout matrix(ncell_3d, ndf, ndf) (Fortran order)."""
    A = np.zeros(ndf * ndf * ncell_3d, np.float64)
    lib().orc_gen_spd_shift(ndf, ncell_3d, seed, int(scaled), float(shift), _p(A))
    return A.reshape((ndf, ndf, ncell_3d)).transpose(2, 1, 0)  # view as (ncell_3d, ndf, ndf)


def gen_op_uniform(n1, n2, ncell_3d, seed, stream, lo, hi, sign=None):
    A = np.zeros(n1 * n2 * ncell_3d, np.float64)
    sg = None if sign is None else np.ascontiguousarray(sign, np.float64)
    lib().orc_gen_op_uniform(n1, n2, ncell_3d, seed, stream, float(lo), float(hi), _p(sg), _p(A))
    return A.reshape((n2, n1, ncell_3d)).transpose(2, 1, 0)


def flat_matrix(A):
    """Fortran-order flat buffer of a (ncell_3d, ndf1, ndf2) matrix view."""
    return np.ascontiguousarray(np.asarray(A).transpose(2, 1, 0)).ravel()


def gen_uniform(n, seed, stream):
    v = np.zeros(n, np.float64)
    lib().orc_gen_uniform(n, seed, stream, _p(v))
    return v


def lma_apply(mesh, A, x, to_mesh=None, r32=False):
    to = to_mesh or mesh
    Af = flat_matrix(A)
    if r32:
        Af = Af.astype(np.float32)
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(to.undf, np.float32)
        fn = lib().orc_lma_apply_r32
    else:
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(to.undf, np.float64)
        fn = lib().orc_lma_apply
    fn(mesh.ncells, mesh.nlayers, len(to.ncells_per_colour), _p(to.ncells_per_colour), _p(to.cells_in_colour),
       to.cells_in_colour.shape[1], to.ndf, to.undf, _p(to.dofmap), mesh.ndf, mesh.undf,
       _p(mesh.dofmap), _p(Af), _p(x), _p(y))
    return y


def mm_diagonal(mesh, A):
    d = np.zeros(mesh.undf, np.float64)
    lib().orc_mm_diagonal(mesh.ncells, mesh.nlayers, mesh.ndf, mesh.undf, _p(mesh.dofmap),
                          _p(flat_matrix(A)), _p(d))
    return d


def invert_local(A, r32=False):
    """Per-cell inverse of a (ncell_3d, ndf, ndf) operator (invert_local_operator_code)."""
    A = np.asarray(A)
    nc3, n, n2 = A.shape
    assert n == n2
    dt = np.float32 if r32 else np.float64
    Af = np.ascontiguousarray(A.astype(dt).transpose(2, 1, 0)).ravel()
    out = np.zeros_like(Af)
    fn = lib().orc_invert_local_r32 if r32 else lib().orc_invert_local
    if fn(nc3, n, _p(Af), _p(out)) != 0:
        raise ValueError("invert_local: ndf out of range")
    return out.reshape((n, n, nc3)).transpose(2, 1, 0)


def gauss_rule(n):
    p, w = np.zeros(n), np.zeros(n)
    if lib().orc_gauss_rule(n, _p(p), _p(w)) != 0:
        raise ValueError("gauss_rule: n out of range")
    return p, w


def mass_matrix_w0(ncells, nlayers, dofmap, chi1, chi2, chi3, nqp=3):
    """Lowest-order W0 mass matrices (ncells*nlayers, 8, 8) from Cartesian chi fields."""
    dm = np.ascontiguousarray(dofmap, np.int32)
    c = [np.ascontiguousarray(v, np.float64) for v in (chi1, chi2, chi3)]
    out = np.zeros(ncells * nlayers * 64)
    if lib().orc_mass_matrix_w0(ncells, nlayers, _p(dm), _p(c[0]), _p(c[1]), _p(c[2]), nqp, _p(out)) != 0:
        raise ValueError("mass_matrix_w0: nqp out of range")
    return out.reshape((8, 8, ncells * nlayers)).transpose(2, 1, 0)


def _chi_args(dofmap, chi1, chi2, chi3):
    dm = np.ascontiguousarray(dofmap, np.int32)
    c = [np.ascontiguousarray(v, np.float64) for v in (chi1, chi2, chi3)]
    return dm, c


def w2_basis(nqp):
    """Lowest-order W2 basis (3, 6, nqp_h, nqp_v) and divergence (1, 6, nqp_h, nqp_v) tables."""
    nh = nqp * nqp
    b, d = np.zeros(3 * 6 * nh * nqp), np.zeros(6 * nh * nqp)
    if lib().orc_w2_basis(nqp, _p(b), _p(d)) != 0:
        raise ValueError("w2_basis: nqp out of range")
    return b.reshape((3, 6, nh, nqp), order="F"), d.reshape((1, 6, nh, nqp), order="F")


def mass_matrix_w2(ncells, nlayers, dofmap, chi1, chi2, chi3, nqp=3):
    """Lowest-order W2 mass matrices (ncells*nlayers, 6, 6); dofmap is the coordinate (W0) map."""
    dm, c = _chi_args(dofmap, chi1, chi2, chi3)
    out = np.zeros(ncells * nlayers * 36)
    if lib().orc_mass_matrix_w2(ncells, nlayers, _p(dm), _p(c[0]), _p(c[1]), _p(c[2]), nqp, _p(out)) != 0:
        raise ValueError("mass_matrix_w2: nqp out of range")
    return out.reshape((6, 6, ncells * nlayers)).transpose(2, 1, 0)


def mass_matrix_w3(ncells, nlayers, dofmap, chi1, chi2, chi3, nqp=3, rehabilitate=True):
    """Lowest-order W3 mass matrices (ncells*nlayers, 1, 1)."""
    dm, c = _chi_args(dofmap, chi1, chi2, chi3)
    out = np.zeros(ncells * nlayers)
    if lib().orc_mass_matrix_w3(ncells, nlayers, _p(dm), _p(c[0]), _p(c[1]), _p(c[2]), nqp, int(bool(rehabilitate)),
                                _p(out)) != 0:
        raise ValueError("mass_matrix_w3: nqp out of range")
    return out.reshape((-1, 1, 1))


def mass_matrix_wtheta(ncells, nlayers, dofmap, chi1, chi2, chi3, nqp=3):
    """Lowest-order Wtheta mass matrices (ncells*nlayers, 2, 2)."""
    dm, c = _chi_args(dofmap, chi1, chi2, chi3)
    out = np.zeros(ncells * nlayers * 4)
    if lib().orc_mass_matrix_wtheta(ncells, nlayers, _p(dm), _p(c[0]), _p(c[1]), _p(c[2]), nqp, _p(out)) != 0:
        raise ValueError("mass_matrix_wtheta: nqp out of range")
    return out.reshape((2, 2, ncells * nlayers)).transpose(2, 1, 0)


def div_operator(ncells, nlayers, dofmap, chi1, chi2, chi3, nqp=3, rehabilitate=True):
    """Weak divergence W2 -> W3 element matrices (ncells*nlayers, 1, 6)."""
    dm, c = _chi_args(dofmap, chi1, chi2, chi3)
    out = np.zeros(ncells * nlayers * 6)
    if lib().orc_div_operator(ncells, nlayers, _p(dm), _p(c[0]), _p(c[1]), _p(c[2]), nqp, int(bool(rehabilitate)),
                              _p(out)) != 0:
        raise ValueError("div_operator: nqp out of range")
    return out.reshape((6, 1, ncells * nlayers)).transpose(2, 1, 0)


# ---- columnwise (CMA) operators --------------------------------------------------------------
# A CMA matrix is held as a C-order (ncells, nrow, bandwidth) array, which is the reference's
# Fortran columnwise_matrix(bandwidth, nrow, ncell_2d) byte for byte; p = [alpha, beta, gamma_m,
# gamma_p, bandwidth, nrow, ncol].

def cma_params(nl, ni_to, nf_to, ni_from, nf_from, custom=None):
    """Banded-structure parameters of a CMA operator between horizontally discontinuous spaces
    with ni interior and nf vertical-face dofs per cell; custom = (alpha, beta, gamma_m, gamma_p)."""
    p = np.zeros(7, np.int32)
    c = custom or (0, 0, 0, 0)
    if lib().orc_cma_params(nl, ni_to, nf_to, ni_from, nf_from, int(custom is not None), *c, _p(p)) != 0:
        raise ValueError("cma_params: alpha and beta are both zero")
    return p


def cma_dofmaps(nl, ndf, ni, nf, map_cell1):
    """column_banded_dofmap as a Fortran-order (ndf, nl) flat array and indirection_dofmap (nrow)."""
    m = np.ascontiguousarray(map_cell1, np.int32)
    banded = np.zeros(ndf * nl, np.int32)
    ind = np.zeros(nl * (ndf - nf) + nf, np.int32)
    lib().orc_cma_dofmaps(nl, ndf, ni, nf, _p(m), _p(banded), _p(ind))
    return banded, ind


def _cma_fn(name, r32):
    return getattr(lib(), name + ("_r32" if r32 else ""))


def _cma_t(r32):
    return np.float32 if r32 else np.float64


def cma_asm(ncells, nl, lma, p, banded_to, banded_from, r32=False):
    """columnwise_op_asm of an LMA operator (ncells*nl, ndf_to, ndf_from)."""
    T = _cma_t(r32)
    L = np.ascontiguousarray(flat_matrix(lma), T)
    out = np.zeros((ncells, p[5], p[4]), T)
    _cma_fn("orc_cma_asm", r32)(ncells, nl, lma.shape[1], lma.shape[2], _p(L), _p(p), _p(banded_to), _p(banded_from),
                                _p(out))
    return out


def cma_apply(p, cma, x, lhs, map1, ind_to, map2, ind_from, r32=False):
    """columnwise_op_app: lhs += A x (returns a new array); map1/map2 are (ncells, ndf) dofmaps."""
    T = _cma_t(r32)
    y = np.array(lhs, T)
    m1, m2 = np.ascontiguousarray(map1, np.int32), np.ascontiguousarray(map2, np.int32)
    xx, A = np.ascontiguousarray(x, T), np.ascontiguousarray(cma, T)
    _cma_fn("orc_cma_apply", r32)(A.shape[0], _p(p), _p(A), _p(xx), _p(y), m1.shape[1], _p(m1), _p(ind_to),
                                  m2.shape[1], _p(m2), _p(ind_from))
    return y


def cma_appinv(p, cma, x, map_, ind, r32=False):
    """columnwise_op_appinv: the tridiagonal solve of every column."""
    T = _cma_t(r32)
    y = np.zeros(len(x), T)
    m, xx, A = np.ascontiguousarray(map_, np.int32), np.ascontiguousarray(x, T), np.ascontiguousarray(cma, T)
    _cma_fn("orc_cma_appinv", r32)(A.shape[0], _p(p), _p(A), _p(xx), _p(y), m.shape[1], _p(m), _p(ind))
    return y


def cma_mul(pa, A, pb, B, pc, C, r32=False):
    """columnwise_op_mul: C + A B (returns a new array)."""
    T = _cma_t(r32)
    out = np.array(C, T)
    Aa, Bb = np.ascontiguousarray(A, T), np.ascontiguousarray(B, T)
    _cma_fn("orc_cma_mul", r32)(Aa.shape[0], _p(pa), _p(Aa), _p(pb), _p(Bb), _p(pc), _p(out))
    return out


def cma_scaledadd(pa, A, pb, B, pc, alpha, beta, r32=False):
    """columnwise_op_scaledadd: alpha A + beta B in C's band structure."""
    T = _cma_t(r32)
    Aa, Bb = np.ascontiguousarray(A, T), np.ascontiguousarray(B, T)
    out = np.zeros((Aa.shape[0], pc[5], pc[4]), T)
    _cma_fn("orc_cma_scaledadd", r32)(Aa.shape[0], _p(pa), _p(Aa), _p(pb), _p(Bb), _p(pc), _p(out), alpha, beta)
    return out


def cma_asm_diag_hmht(ncells, nl, Dh, M2h, p, banded, r32=False):
    """columnwise_op_asm_diag_hmht from LMA Dh (c3, ndf_w3, ndf_w2h) and M2h (c3, ndf_w2h, ndf_w2h)."""
    T = _cma_t(r32)
    D, M = np.ascontiguousarray(flat_matrix(Dh), T), np.ascontiguousarray(flat_matrix(M2h), T)
    out = np.zeros((ncells, p[5], p[4]), T)
    _cma_fn("orc_cma_asm_diag_hmht", r32)(ncells, nl, Dh.shape[1], Dh.shape[2], _p(D), _p(M), _p(p), _p(banded),
                                          _p(out))
    return out


def cma_asm_m2v_lumped_inv(ncells, nl, lma, p, banded, r32=False):
    """columnwise_op_asm_m2v_lumped_inv of a square LMA operator."""
    T = _cma_t(r32)
    L = np.ascontiguousarray(flat_matrix(lma), T)
    out = np.zeros((ncells, p[5], p[4]), T)
    _cma_fn("orc_cma_asm_m2v_lumped_inv", r32)(ncells, nl, lma.shape[1], _p(L), _p(p), _p(banded), _p(out))
    return out


def tri_solve(nl, map_, x, t0, tp, tm, r32=False):
    """tri_solve_code over every W3 column (map: the W3 dofmap, one dof per cell)."""
    T = _cma_t(r32)
    m = np.ascontiguousarray(np.asarray(map_).ravel(), np.int32)
    a = [np.ascontiguousarray(v, T) for v in (x, t0, tp, tm)]
    y = np.zeros(len(x), T)
    _cma_fn("orc_tri_solve", r32)(len(m), nl, _p(m), *[_p(v) for v in a], _p(y))
    return y


def make_problem(mesh, A, diag=None):
    Af = flat_matrix(A)
    keep = [Af, diag]
    prob = OrcProblem(mesh.ncells, mesh.nlayers, len(mesh.ncells_per_colour), mesh.cells_in_colour.shape[1], mesh.ndf,
                      mesh.undf, _p(mesh.dofmap).value, _p(mesh.ncells_per_colour).value,
                      _p(mesh.cells_in_colour).value, _p(Af).value,
                      _p(diag).value if diag is not None else None, None, None, None)
    prob._keep = keep
    return prob


def cg(mesh, A, b, x0=None, diag=None, r_tol=1e-8, a_tol=1e-16, max_iter=500, monitor=True):
    prob = make_problem(mesh, A, diag)
    x = np.zeros(mesh.undf) if x0 is None else np.array(x0, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    res = OrcResult()
    hist = np.zeros(max_iter + 1)
    lib().orc_cg(ctypes.byref(prob), _p(b), _p(x), r_tol, a_tol, max_iter, int(monitor), ctypes.byref(res), _p(hist))
    return x, res, hist[: res.iterations + 1]


def cg_r32(mesh, A, b, diag, r_tol=1e-8, a_tol=0.0, max_iter=500):
    """cg_solve (components/science/source/solver/sci_iterative_solver_mod.f90:495-592, monitor on)
    over r_solver vectors (sci_r_solver_field_vector_mod.x90:119-252): real32 storage, the
    real(r_def) scalars rounded to real32 before use (:189,213-214,246), inner products of the
    real32 entries accumulated in real64 (sci_psykal_light_mod.f90:72-214), the real32 LMA apply
    (matrix_vector_code_r_single) and z = r / diag in real32.  axpy and aypx round once per
    entry (a fused multiply-add: the product of two real32 values is exact in real64).
    Returns (x, iterations, converged)."""
    f32, f64 = np.float32, np.float64
    d = np.asarray(diag, f32)
    b = np.asarray(b, f32)

    def apply(v):
        return lma_apply(mesh, A, v, r32=True)

    def dot(u, v):
        return float(np.dot(u.astype(f64), v.astype(f64)))

    def fma(a, u, v):  # v + real(a, r_solver) * u, rounded once
        return (f64(f32(a)) * u.astype(f64) + v.astype(f64)).astype(f32)

    x = np.zeros(mesh.undf, f32)
    r = apply(x)
    r = (f32(-1.0) * r).astype(f32)
    r = fma(1.0, b, r)
    r0 = np.sqrt(dot(r, r))
    if r0 < np.finfo(f64).eps:
        return x, 0, True
    z = (r / d).astype(f32)
    rz = dot(r, z)
    p = z.copy()
    for it in range(1, max_iter + 1):
        z = apply(p)
        alpha = rz / dot(p, z)
        x = fma(alpha, p, x)
        r = fma(-alpha, z, r)
        rn = np.sqrt(dot(r, r))
        if rn / r0 <= r_tol or rn <= a_tol:
            return x, it, True
        z = (r / d).astype(f32)
        rz_new = dot(r, z)
        beta = rz_new / rz
        p = (f64(f32(beta)) * p.astype(f64) + z.astype(f64)).astype(f32)
        rz = rz_new
    return x, max_iter, False


def gmres_callbacks(n, apply, prec, b, x0=None, gcrk=30, r_tol=1e-8, a_tol=1e-16, max_iter=50, monitor=True):
    """GMRES(gcrk) of the oracle with python-callable operator and preconditioner."""

    def wrap(f):
        def cb(ctx, xp, yp):
            xv = np.ctypeslib.as_array(xp, shape=(n,))
            yv = np.ctypeslib.as_array(yp, shape=(n,))
            yv[:] = f(xv.copy())
        return CALLBACK(cb)

    ca, cp = wrap(apply), wrap(prec)
    prob = OrcProblem(0, 0, 0, 0, 0, n, None, None, None, None, None,
                      ctypes.cast(ca, _VP).value, ctypes.cast(cp, _VP).value, None)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    res = OrcResult()
    hist = np.zeros(max_iter + 1)
    lib().orc_gmres(ctypes.byref(prob), _p(b), _p(x), gcrk, r_tol, a_tol, max_iter, int(monitor),
                    ctypes.byref(res), _p(hist))
    return x, res, hist[: res.iterations + 1]


def cg_callbacks(n, apply, prec, b, x0=None, r_tol=1e-8, a_tol=1e-16, max_iter=500, monitor=True):
    def wrap(f):
        def cb(ctx, xp, yp):
            xv = np.ctypeslib.as_array(xp, shape=(n,))
            yv = np.ctypeslib.as_array(yp, shape=(n,))
            yv[:] = f(xv.copy())
        return CALLBACK(cb)

    ca, cp = wrap(apply), wrap(prec)
    prob = OrcProblem(0, 0, 0, 0, 0, n, None, None, None, None, None,
                      ctypes.cast(ca, _VP).value, ctypes.cast(cp, _VP).value, None)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    res = OrcResult()
    hist = np.zeros(max_iter + 1)
    lib().orc_cg(ctypes.byref(prob), _p(b), _p(x), r_tol, a_tol, max_iter, int(monitor), ctypes.byref(res), _p(hist))
    return x, res, hist[: res.iterations + 1]


def givens_lsq(h, g):
    m = h.shape[1]
    hf = np.asfortranarray(h, np.float64).ravel(order="F")
    u = np.zeros(m)
    lib().orc_givens_lsq(m, _p(np.ascontiguousarray(hf)), _p(np.ascontiguousarray(g, np.float64)), _p(u))
    return u


def assembled_csr(mesh, A, to_mesh=None):
    """Global sparse matrix sum_c P_c^T A_c P_c assembled with scipy (brute-force check)."""
    import scipy.sparse as sp
    to = to_mesh or mesh
    A = np.asarray(A)
    nl = mesh.nlayers
    rows, cols, vals = [], [], []
    cells = np.arange(mesh.ncells)
    for k in range(nl):
        ik = cells * nl + k
        for i in range(to.ndf):
            for j in range(mesh.ndf):
                rows.append(to.dofmap[:, i] - 1 + k)
                cols.append(mesh.dofmap[:, j] - 1 + k)
                vals.append(A[ik, i, j])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(to.undf, mesh.undf))


def transpose_apply(to_mesh, from_mesh, A, x):
    """y (on to_mesh) = A^T x, A of shape (ncell_3d, from_ndf, to_ndf) as the reference holds it
    for transpose_matrix_vector_kernel (lhs on ANY_SPACE_1 = to_mesh)."""
    y = np.zeros(to_mesh.undf)
    Af = flat_matrix(A)
    lib().orc_transpose_apply(to_mesh.ncells, to_mesh.nlayers, len(to_mesh.ncells_per_colour), _p(to_mesh.ncells_per_colour),
                              _p(to_mesh.cells_in_colour), to_mesh.cells_in_colour.shape[1], to_mesh.ndf,
                              to_mesh.undf, _p(to_mesh.dofmap), from_mesh.ndf, from_mesh.undf,
                              _p(from_mesh.dofmap), _p(Af), _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y


def dg_apply(to_mesh, from_mesh, A, x):
    y = np.zeros(to_mesh.undf)
    lib().orc_dg_apply(to_mesh.ncells, to_mesh.nlayers, to_mesh.ndf, to_mesh.undf, _p(to_mesh.dofmap),
                       from_mesh.ndf, from_mesh.undf, _p(from_mesh.dofmap), _p(flat_matrix(A)),
                       _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y


def cubedsphere_lonlat(n):
    """(longitude, latitude) in radians of the mesh generator's vertices, by global vertex id."""
    nv = 6 * n * n + 2
    lon, lat = np.zeros(nv), np.zeros(nv)
    if lib().orc_cubedsphere_lonlat(n, _p(lon), _p(lat)) != 0:
        raise ValueError(n)
    return lon, lat


def eta(kind, nl):
    """eta(0:nl) of an extrusion: 0 uniform, 1 quadratic, 2 geometric."""
    e = np.zeros(nl + 1)
    if lib().orc_eta(kind, nl, _p(e)) != 0:
        raise ValueError((kind, nl))
    return e


def chi_w0(mesh, kind, domain_top, scaled_radius):
    """The Cartesian coordinate fields (chi1, chi2, chi3) on a lowest-order W0 cubed-sphere mesh."""
    lon, lat = cubedsphere_lonlat(mesh.n)
    vlon, vlat = np.zeros(mesh.nverts), np.zeros(mesh.nverts)
    g, l = np.asarray(mesh.vert_gid).ravel(), np.asarray(mesh.vert_on_cell).ravel()
    vlon[l - 1], vlat[l - 1] = lon[g - 1], lat[g - 1]
    chi = [np.zeros(mesh.undf) for _ in range(3)]
    voc = np.ascontiguousarray(mesh.vert_on_cell, np.int32)
    if lib().orc_chi_w0(mesh.ncells, mesh.nlayers, _p(mesh.dofmap), _p(voc), _p(vlon), _p(vlat), kind,
                        float(domain_top), float(scaled_radius), _p(chi[0]), _p(chi[1]), _p(chi[2])) != 0:
        raise ValueError(kind)
    return chi


def dg_inc_apply(to_mesh, from_mesh, A, x, y0):
    """y0 + A x on a discontinuous to_mesh (dg_inc_matrix_vector_code); y0 is not modified."""
    y = np.array(y0, np.float64, copy=True)
    lib().orc_dg_inc_apply(to_mesh.ncells, to_mesh.nlayers, to_mesh.ndf, to_mesh.undf, _p(to_mesh.dofmap),
                           from_mesh.ndf, from_mesh.undf, _p(from_mesh.dofmap), _p(flat_matrix(A)),
                           _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y


def boundary_dofs(space, order, ndf):
    """dof_on_vert_boundary (ndf, 2) of a lowest-order space: column 0 bottom, 1 top."""
    bv = np.zeros(2 * ndf, np.int32)
    if lib().orc_boundary_dofs(space, order, ndf, _p(bv)) != 0:
        raise ValueError("boundary dofs not restated for this space")
    return bv.reshape(2, ndf).T.copy()


def enforce_bc(mesh, bv, field):
    """enforce_bc_code over every cell column; bv is (ndf, 2); returns a new field."""
    y = np.array(field, np.float64, copy=True)
    bvf = np.ascontiguousarray(np.asarray(bv, np.int32).T).ravel()  # Fortran order (ndf, 2)
    lib().orc_enforce_bc(mesh.ncells, mesh.nlayers, mesh.ndf, _p(mesh.dofmap), _p(bvf), _p(y))
    return y


D_SIGNS = np.array([-1.0, -1.0, 1.0, 1.0, -1.0, 1.0])  # outward normals of faces W,S,E,N,B,T


class Helmholtz:
    """H = M3 + tau D diag(M2)^-1 D^T on W3 with the synthetic element matrices of the mixed
    W2/W3 configuration: M2_e = C C^T + I (6x6, C ~ N(0,1)/8), D_e = sign(face) U(0.5,1.5)
    (1x6), M3_e = c^2 + 1 (c ~ N(0,1)/8)."""

    def __init__(self, n, nl, seed=0, tau=0.5):
        self.w2 = Mesh(n, nl, W2, 0)
        self.w3 = Mesh(n, nl, W3, 0)
        nc3 = self.w3.ncell_3d
        self.M2 = gen_spd(6, nc3, seed, shift=1.0)
        self.D = gen_op_uniform(1, 6, nc3, seed, 5, 0.5, 1.5, D_SIGNS)
        self.M3 = gen_spd(1, nc3, seed + 1, shift=1.0)
        self.m2d = mm_diagonal(self.w2, self.M2)
        self.tau = tau

    def apply(self, x):
        t = transpose_apply(self.w2, self.w3, self.D, x) / self.m2d
        return dg_apply(self.w3, self.w3, self.M3, x) + self.tau * dg_apply(self.w3, self.w2, self.D, t)

    def dense(self):
        n = self.w3.undf
        return np.column_stack([self.apply(np.eye(1, n, i).ravel()) for i in range(n)])


METHODS = ("cg", "bicgstab", "gmres", "fgmres", "gcr", "prec_only", "jacobi", "chebyshev")


def solve_callbacks(method, n, apply, prec, b, x0=None, gcrk=30, r_tol=1e-8, a_tol=1e-16, max_iter=500,
                    monitor=True, rho_relax=1.0, lmin=0.124, lmax=3.377):
    """Any solver of the mass-matrix solver factory (sci_mass_matrix_solver_alg_mod.x90:164-181)
    with python-callable operator and preconditioner.  Returns (x, result, history)."""

    def wrap(f):
        def cb(ctx, xp, yp):
            xv = np.ctypeslib.as_array(xp, shape=(n,))
            yv = np.ctypeslib.as_array(yp, shape=(n,))
            yv[:] = f(xv.copy())
        return CALLBACK(cb)

    ca, cp = wrap(apply), wrap(prec)
    prob = OrcProblem(0, 0, 0, 0, 0, n, None, None, None, None, None,
                      ctypes.cast(ca, _VP).value, ctypes.cast(cp, _VP).value, None)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    res = OrcResult()
    hist = np.zeros(max_iter * max(gcrk, 1) + 2)
    P, R, L = ctypes.byref(prob), ctypes.byref(res), lib()
    if method == "cg":
        L.orc_cg(P, _p(b), _p(x), r_tol, a_tol, max_iter, int(monitor), R, _p(hist))
    elif method == "bicgstab":
        L.orc_bicgstab(P, _p(b), _p(x), r_tol, a_tol, max_iter, int(monitor), R, _p(hist))
    elif method == "gmres":
        L.orc_gmres(P, _p(b), _p(x), gcrk, r_tol, a_tol, max_iter, int(monitor), R, _p(hist))
    elif method == "fgmres":
        L.orc_fgmres(P, _p(b), _p(x), gcrk, r_tol, a_tol, max_iter, int(monitor), R, _p(hist))
    elif method == "gcr":
        L.orc_gcr(P, _p(b), _p(x), gcrk, r_tol, a_tol, max_iter, int(monitor), R, _p(hist))
    elif method == "jacobi":
        L.orc_jacobi(P, _p(b), _p(x), r_tol, a_tol, max_iter, int(monitor), rho_relax, R, _p(hist))
    elif method == "chebyshev":
        L.orc_chebyshev(P, _p(b), _p(x), max_iter, lmin, lmax, R)
    elif method == "prec_only":
        L.orc_prec_only(P, _p(b), _p(x), R)
    else:
        raise ValueError(method)
    return x, res, hist
