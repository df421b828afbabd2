"""Function spaces, field vectors and LMA operators resident on the GPU."""
import ctypes

import numpy as np

from ._lib import R32, R64, W0, check, init, lib, ptr


class FunctionSpace:
    """Mesh + function space inputs of the hot path: the column-contiguous dofmap
    (infrastructure/source/function_space/function_space_mod.F90:733-757) and the colour map
    (infrastructure/source/mesh/mesh_mod.F90:2202-2245), held on the device."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle) if not isinstance(handle, ctypes.c_void_p) else handle
        n = [ctypes.c_int() for _ in range(5)]
        check(lib().lfric_fs_info(self._h, *[ctypes.byref(v) for v in n]))
        self.ncells, self.nlayers, self.ndf, self.undf, self.ncolours = [v.value for v in n]

    @classmethod
    def cubedsphere(cls, edge_cells, nlayers, space=W0, order=0):
        init()
        h = ctypes.c_void_p()
        check(lib().lfric_fs_create_cubedsphere(edge_cells, nlayers, space, order, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_dofmap(cls, nlayers, undf, dofmap, ncells_per_colour, cells_in_colour):
        """dofmap (ncells, ndf) 1-based; cells_in_colour (ncolours, maxcc) 1-based."""
        init()
        dm = np.ascontiguousarray(dofmap, np.int32)
        npc = np.ascontiguousarray(ncells_per_colour, np.int32)
        cic = np.ascontiguousarray(cells_in_colour, np.int32)
        h = ctypes.c_void_p()
        check(lib().lfric_fs_create(dm.shape[0], nlayers, dm.shape[1], undf, ptr(dm), cic.shape[0], ptr(npc),
                                    ptr(cic), cic.shape[1], ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_ugrid(cls, path, nlayers, space=W0, order=0, mesh=None):
        """A space on the extruded global mesh of a UGRID 2D file (ugrid_2d_mod.F90's
        face_node and face_face connectivity), coloured as the reference colours that mesh."""
        init()
        h = ctypes.c_void_p()
        check(lib().lfric_fs_create_ugrid(str(path).encode(), mesh.encode() if mesh else None, nlayers, space, order,
                                          ctypes.byref(h)))
        return cls(h)

    def coordinates(self, extrusion=1, domain_top=80000.0, scaled_radius=6371229.0):
        """The Cartesian coordinate fields (chi1, chi2, chi3) of a generated lowest-order W0
        cubed-sphere space: the mesh generator's vertices on spheres of radius scaled_radius +
        height, heights from the extrusion (0 uniform, 1 quadratic, 2 geometric) to domain_top
        (assign_coordinate_field, components/driver/source/driver_coordinates_mod.F90:59-454)."""
        chi = [FieldVector(self) for _ in range(3)]
        check(lib().lfric_fs_coordinates(self._h, int(extrusion), float(domain_top), float(scaled_radius),
                                         chi[0]._h, chi[1]._h, chi[2]._h))
        return chi

    def boundary_dofs(self):
        """get_boundary_dofs (function_space_mod.F90:791-801): (ndf, 2), column 0 bottom,
        column 1 top; 0 flags a dof on that boundary."""
        bv = np.zeros(2 * self.ndf, np.int32)
        check(lib().lfric_fs_get_boundary_dofs(self._h, ptr(bv)))
        return bv.reshape(2, self.ndf).T.copy()

    def set_boundary_dofs(self, bv):
        a = np.ascontiguousarray(np.asarray(bv, np.int32).T).ravel()
        assert a.size == 2 * self.ndf
        check(lib().lfric_fs_set_boundary_dofs(self._h, ptr(a)))

    def sibling(self, space, order=0):
        h = ctypes.c_void_p()
        check(lib().lfric_fs_create_sibling(self._h, space, order, ctypes.byref(h)))
        return FunctionSpace(h)

    @property
    def ncell_3d(self):
        return self.ncells * self.nlayers

    def dofmap(self):
        dm = np.zeros((self.ncells, self.ndf), np.int32)
        check(lib().lfric_fs_get_dofmap(self._h, ptr(dm)))
        return dm

    def colouring(self):
        npc = np.zeros(self.ncolours, np.int32)
        check(lib().lfric_fs_get_colouring(self._h, ptr(npc), None))
        cic = np.zeros((self.ncolours, int(npc.max())), np.int32)
        check(lib().lfric_fs_get_colouring(self._h, ptr(npc), ptr(cic)))
        return npc, cic

    def __del__(self):
        if getattr(self, "_h", None) and lib is not None:
            try:
                lib().lfric_fs_destroy(self._h)
            except Exception:
                pass
            self._h = None


class FieldVector:
    """abstract_vector_type on the device (infrastructure/source/field/vector_mod.F90:17-59).
    precision R64 mirrors field_vector_type, R32 mirrors r_solver_field_vector_type (FP32
    data, FP64-accumulated dot and norm)."""

    def __init__(self, fs, precision=R64, nrhs=1):
        self.fs = fs
        self.precision = precision
        self.nrhs = nrhs
        self._h = ctypes.c_void_p()
        check(lib().lfric_vec_create(fs._h, precision, nrhs, ctypes.byref(self._h)))

    # import_field / export_field (sci_field_vector_mod.x90:294-334)
    def import_field(self, data):
        a = np.ascontiguousarray(data, np.float64).ravel()
        assert a.size == self.fs.undf * self.nrhs
        check(lib().lfric_vec_upload(self._h, ptr(a)))
        return self

    def export_field(self):
        out = np.zeros(self.fs.undf * self.nrhs, np.float64)
        check(lib().lfric_vec_download(self._h, ptr(out)))
        return out

    def export_raw(self):
        if self.precision == R32:
            out = np.zeros(self.fs.undf * self.nrhs, np.float32)
            check(lib().lfric_vec_download_r32(self._h, ptr(out)))
            return out
        return self.export_field()

    def import_raw(self, data):
        if self.precision == R32:
            a = np.ascontiguousarray(data, np.float32).ravel()
            check(lib().lfric_vec_upload_r32(self._h, ptr(a)))
            return self
        return self.import_field(data)

    def set_scalar(self, scalar):
        check(lib().lfric_vec_set_scalar(self._h, float(scalar)))

    def axpy(self, alpha, x):
        check(lib().lfric_vec_axpy(self._h, float(alpha), x._h))

    def aypx(self, alpha, x):
        check(lib().lfric_vec_aypx(self._h, float(alpha), x._h))

    def axpby(self, a, x, b, y):
        check(lib().lfric_vec_axpby(self._h, float(a), x._h, float(b), y._h))

    def scale(self, alpha):
        check(lib().lfric_vec_scale(self._h, float(alpha)))

    def copy(self, source):
        check(lib().lfric_vec_copy(self._h, source._h))

    def enforce_bc(self):
        """enforce_bc_kernel_type on this field (sci_enforce_bc_kernel_mod.F90:63-91)."""
        check(lib().lfric_enforce_bc(self._h))
        return self

    def divide(self, x, d):
        check(lib().lfric_vec_divide(self._h, x._h, d._h))

    def dot(self, x):
        out = ctypes.c_double()
        check(lib().lfric_vec_dot(self._h, x._h, ctypes.byref(out)))
        return out.value

    def norm(self):
        out = ctypes.c_double()
        check(lib().lfric_vec_norm(self._h, ctypes.byref(out)))
        return out.value

    def duplicate(self):
        return FieldVector(self.fs, self.precision, self.nrhs)

    def vector_size(self):
        return self.fs.undf * self.nrhs

    def fill_uniform(self, seed, stream=3):
        check(lib().lfric_vec_fill_uniform(self._h, seed, stream))
        return self

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().lfric_vec_destroy(self._h)
            except Exception:
                pass
            self._h = None


class LmaOperator:
    """operator_type (infrastructure/source/operator/operator_mod.t90:31-137): one dense
    ndf_to x ndf_from element matrix per 3D cell, held on the device; apply() is the invoke of
    setval_c + matrix_vector_kernel_type (sci_mass_matrix_operator_alg_mod.x90:98-100)."""

    def __init__(self, fs_to, fs_from=None, precision=R64):
        self.fs_to = fs_to
        self.fs_from = fs_from or fs_to
        self.precision = precision
        self._h = ctypes.c_void_p()
        check(lib().lfric_op_create(self.fs_to._h, self.fs_from._h, precision, ctypes.byref(self._h)))

    @property
    def shape(self):
        return (self.fs_to.ncell_3d, self.fs_to.ndf, self.fs_from.ndf)

    def upload(self, local_stencil):
        """local_stencil (ncell_3d, ndf_to, ndf_from) as the reference holds it."""
        a = np.asarray(local_stencil, np.float64)
        assert a.shape == self.shape
        flat = np.ascontiguousarray(a.transpose(2, 1, 0)).ravel()  # Fortran order
        check(lib().lfric_op_upload(self._h, ptr(flat)))
        return self

    def download(self):
        n1, n2 = self.fs_to.ndf, self.fs_from.ndf
        flat = np.zeros(self.fs_to.ncell_3d * n1 * n2, np.float64)
        check(lib().lfric_op_download(self._h, ptr(flat)))
        return flat.reshape((n2, n1, self.fs_to.ncell_3d)).transpose(2, 1, 0)

    def generate_spd(self, seed, scaled=False, shift=0.1):
        check(lib().lfric_op_generate_spd(self._h, seed, int(scaled), float(shift)))
        return self

    def generate_uniform(self, seed, stream, lo, hi, col_sign=None):
        s = None if col_sign is None else ptr(np.ascontiguousarray(col_sign, np.float64))
        check(lib().lfric_op_generate_uniform(self._h, seed, stream, float(lo), float(hi), s))
        return self

    def convert(self, precision):
        out = LmaOperator(self.fs_to, self.fs_from, precision)
        check(lib().lfric_op_convert(self._h, out._h))
        return out

    def assemble_mass_w0(self, chi1, chi2, chi3, nqp=3):
        """Every cell's lowest-order W0 mass matrix from the Cartesian coordinate fields chi1..3
        (compute_mass_matrix_w_scalar_code, components/science/source/kernel/fem/
        sci_compute_mass_matrix_kernel_w_scalar_mod.F90:290-384); FP64 operator."""
        check(lib().lfric_op_assemble_mass_w0(self._h, chi1._h, chi2._h, chi3._h, int(nqp)))
        return self

    def assemble_mass(self, chi1, chi2, chi3, nqp=3, rehabilitate=True):
        """Every cell's lowest-order mass matrix of the operator's space (W0, W2 or W3, told
        from its dofmap) from Cartesian coordinate fields on a W0 space of the same mesh:
        compute_mass_matrix_w_scalar_code / compute_mass_matrix_w2_code /
        compute_mass_matrix_w3_code (components/science/source/kernel/fem/
        sci_compute_mass_matrix_kernel_{w_scalar,w2,w3}_mod.F90); FP64 operator."""
        check(lib().lfric_op_assemble_mass(self._h, chi1._h, chi2._h, chi3._h, int(nqp), int(bool(rehabilitate))))
        return self

    def assemble_div(self, chi1, chi2, chi3, nqp=3, rehabilitate=True):
        """The weak divergence W2 -> W3 of every cell (compute_div_operator_code,
        components/science/source/kernel/fem/sci_compute_div_operator_kernel_mod.F90:120-150);
        FP64 operator from a lowest-order W2 space to a W3 space."""
        check(lib().lfric_op_assemble_div(self._h, chi1._h, chi2._h, chi3._h, int(nqp), int(bool(rehabilitate))))
        return self

    def invert_local(self):
        """Per-cell inverse of every element matrix (invert_local_operator_code,
        components/science/source/kernel/algebra/sci_invert_local_operator_kernel_mod.F90:51-77),
        as a new operator on the same space."""
        out = LmaOperator(self.fs_to, self.fs_from, self.precision)
        check(lib().lfric_op_invert_local(self._h, out._h))
        return out

    def apply(self, x, y):
        """y = A x (abstract_linear_operator_type%apply)."""
        check(lib().lfric_lma_apply(self._h, x._h, y._h))

    def tune(self, form=1):
        """Autotune the apply's XCD block order for this operator; returns (window, ms)."""
        w, ms = ctypes.c_int(), ctypes.c_double()
        check(lib().lfric_op_tune(self._h, int(form), ctypes.byref(w), ctypes.byref(ms)))
        return w.value, ms.value

    def prepare_rows(self):
        """Row-gathered copy for the gather apply (square lowest-order W0 operators)."""
        check(lib().lfric_op_prepare_rows(self._h))
        return self

    def prepare_symmetric(self, levels_per_thread=0):
        """Symmetric-packed copy (square W0 operators of order 0 or 1 with symmetric element
        matrices): the gather apply (one or several right-hand sides, the solvers' apply and the
        batch's) then reads the upper triangles only, with the colour-ordered apply's result bit
        for bit.  Order 0: levels_per_thread 0 is one (cell, level) unit per thread, the previous
        layer's top rows passed between lanes; S > 0: S units per thread, each thread's copy
        repeating the previous segment's last triangle.  Order 1: one (cell, layer) unit per
        thread (levels_per_thread 0), its 18 outputs kept between the two phases."""
        check(lib().lfric_op_prepare_symmetric(self._h, int(levels_per_thread)))
        return self

    def time_apply(self, x, y, reps=20):
        """Average ms of the gather apply's kernels over `reps` applies (HIP events): (phase 1,
        phase 2) of the symmetric-packed form, or (the one kernel, 0)."""
        out = np.zeros(2, np.float64)
        check(lib().lfric_op_time_apply(self._h, x._h, y._h, int(reps),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return float(out[0]), float(out[1])

    def set_window(self, window):
        check(lib().lfric_op_set_window(self._h, int(window)))
        return self

    def apply_mrhs(self, x, y, variant=0):
        """y = A x for every right-hand side of x (x.nrhs == y.nrhs); variant 0 auto,
        1 gather form, 2 colour-ordered scatter, 3 colour-ordered scatter on the matrix cores
        (square lowest-order W0 FP32 operators; bit-identical to 2)."""
        check(lib().lfric_lma_apply_mrhs(self._h, x._h, y._h, int(variant)))
        return y

    def apply_gather(self, x, y):
        check(lib().lfric_lma_apply_gather(self._h, x._h, y._h))

    def apply_transpose(self, x, y):
        check(lib().lfric_lma_apply_transpose(self._h, x._h, y._h))

    def apply_dg_inc(self, x, y):
        """y = y + A x (dg_inc_matrix_vector_kernel_mod.F90:69-142)."""
        check(lib().lfric_dg_inc_apply(self._h, x._h, y._h))

    def apply_dg(self, x, y):
        check(lib().lfric_dg_apply(self._h, x._h, y._h))

    def mm_diagonal(self, diag):
        check(lib().lfric_mm_diagonal(self._h, diag._h))
        return diag

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().lfric_op_destroy(self._h)
            except Exception:
                pass
            self._h = None


def extrusion_eta(kind, nlayers):
    """eta(0:nlayers) of an extrusion: 0 uniform, 1 quadratic, 2 geometric (extrusion_mod.f90)."""
    e = np.zeros(nlayers + 1)
    check(lib().lfric_extrusion_eta(int(kind), int(nlayers), ptr(e)))
    return e


def read_ugrid(path, mesh=None):
    """(face_nodes, face_links, npanels) of a UGRID 2D file: 1-based (nfaces, 4) int32 arrays,
    0 for a missing neighbour.  Host only."""
    n, npan = ctypes.c_int(), ctypes.c_int()
    m = mesh.encode() if mesh else None
    check(lib().lfric_ugrid_read(str(path).encode(), m, ctypes.byref(n), ctypes.byref(npan), None, None))
    fn = np.zeros((n.value, 4), np.int32)
    fl = np.zeros((n.value, 4), np.int32)
    check(lib().lfric_ugrid_read(str(path).encode(), m, ctypes.byref(n), ctypes.byref(npan), ptr(fn), ptr(fl)))
    return fn, fl, npan.value
