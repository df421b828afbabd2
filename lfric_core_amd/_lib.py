"""ctypes binding of liblfric_hip.so (the C ABI declared in include/lfric_hip.h)."""
import ctypes
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("LFRIC_HIP_LIB", PKG / "liblfric_hip.so"))

LFRIC_OK = 0
LFRIC_NOT_CONVERGED = 1
LFRIC_NAN = 2
LFRIC_BREAKDOWN = 3
R64, R32, BF16, F16 = 0, 1, 2, 3
W0, W2, W3 = 0, 2, 3

STATUS_NAMES = {0: "ok", 1: "not converged", 2: "NaN", 3: "breakdown", -1: "invalid argument",
                -2: "HIP error", -3: "out of device memory", -4: "bad state", -5: "unsupported"}


class LfricError(RuntimeError):
    """A failed library call (the reference aborts through log_event at LOG_LEVEL_ERROR)."""

    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class NotConverged(LfricError):
    pass


_VP = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_ULL = ctypes.c_ulonglong
_DP = ctypes.POINTER(ctypes.c_double)


class SolverOpts(ctypes.Structure):
    _fields_ = [("r_tol", _D), ("a_tol", _D), ("max_iter", _I), ("gcrk", _I),
                ("monitor_convergence", _I), ("fail_on_non_converged", _I), ("check_every", _I),
                ("inner_r_tol", _D), ("max_outer", _I), ("rho_relax", _D), ("lmin", _D), ("lmax", _D)]


class SolverRes(ctypes.Structure):
    _fields_ = [("status", _I), ("iterations", _I), ("outer_iterations", _I), ("r_nrm_0", _D),
                ("r_nrm", _D), ("true_rel_res", _D)]


# name -> (restype, argtypes); the exported-symbol test checks this table against the header
SIGNATURES = {
    "lfric_init": (_I, [_I]),
    "lfric_last_error": (ctypes.c_char_p, []),
    "lfric_synchronize": (_I, []),
    "lfric_timer_start": (_I, []),
    "lfric_timer_stop": (_I, [_DP]),
    "lfric_device_bytes": (ctypes.c_longlong, []),
    "lfric_set_apply_form": (_I, [_I]),
    "lfric_device_memory": (_I, [ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    "lfric_fs_create": (_I, [_I, _I, _I, _I, _VP, _I, _VP, _VP, _I, ctypes.POINTER(_VP)]),
    "lfric_fs_create_cubedsphere": (_I, [_I, _I, _I, _I, ctypes.POINTER(_VP)]),
    "lfric_fs_create_sibling": (_I, [_VP, _I, _I, ctypes.POINTER(_VP)]),
    "lfric_fs_info": (_I, [_VP] + [ctypes.POINTER(_I)] * 5),
    "lfric_fs_get_dofmap": (_I, [_VP, _VP]),
    "lfric_fs_get_colouring": (_I, [_VP, _VP, _VP]),
    "lfric_fs_destroy": (None, [_VP]),
    "lfric_vec_create": (_I, [_VP, _I, _I, ctypes.POINTER(_VP)]),
    "lfric_vec_upload": (_I, [_VP, _VP]),
    "lfric_vec_download": (_I, [_VP, _VP]),
    "lfric_vec_upload_r32": (_I, [_VP, _VP]),
    "lfric_vec_download_r32": (_I, [_VP, _VP]),
    "lfric_vec_size": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "lfric_vec_destroy": (None, [_VP]),
    "lfric_vec_set_scalar": (_I, [_VP, _D]),
    "lfric_vec_axpy": (_I, [_VP, _D, _VP]),
    "lfric_vec_aypx": (_I, [_VP, _D, _VP]),
    "lfric_vec_axpby": (_I, [_VP, _D, _VP, _D, _VP]),
    "lfric_vec_scale": (_I, [_VP, _D]),
    "lfric_vec_copy": (_I, [_VP, _VP]),
    "lfric_vec_divide": (_I, [_VP, _VP, _VP]),
    "lfric_vec_dot": (_I, [_VP, _VP, _DP]),
    "lfric_vec_norm": (_I, [_VP, _DP]),
    "lfric_vec_fill_uniform": (_I, [_VP, _ULL, _ULL]),
    "lfric_op_create": (_I, [_VP, _VP, _I, ctypes.POINTER(_VP)]),
    "lfric_op_upload": (_I, [_VP, _VP]),
    "lfric_op_download": (_I, [_VP, _VP]),
    "lfric_op_upload_r32": (_I, [_VP, _VP]),
    "lfric_op_download_r32": (_I, [_VP, _VP]),
    "lfric_op_generate_spd": (_I, [_VP, _ULL, _I, _D]),
    "lfric_op_generate_uniform": (_I, [_VP, _ULL, _ULL, _D, _D, _VP]),
    "lfric_op_convert": (_I, [_VP, _VP]),
    "lfric_op_invert_local": (_I, [_VP, _VP]),
    "lfric_op_assemble_mass_w0": (_I, [_VP, _VP, _VP, _VP, _I]),
    "lfric_op_assemble_mass": (_I, [_VP, _VP, _VP, _VP, _I, _I]),
    "lfric_op_assemble_div": (_I, [_VP, _VP, _VP, _VP, _I, _I]),
    "lfric_op_prepare_symmetric": (_I, [_VP, _I]),
    "lfric_op_destroy": (None, [_VP]),
    "lfric_lma_apply": (_I, [_VP, _VP, _VP]),
    "lfric_lma_apply_gather": (_I, [_VP, _VP, _VP]),
    "lfric_lma_apply_transpose": (_I, [_VP, _VP, _VP]),
    "lfric_dg_apply": (_I, [_VP, _VP, _VP]),
    "lfric_mm_diagonal": (_I, [_VP, _VP]),
    "lfric_prec_create_diagonal": (_I, [_VP, ctypes.POINTER(_VP)]),
    "lfric_prec_create_identity": (_I, [ctypes.POINTER(_VP)]),
    "lfric_prec_apply": (_I, [_VP, _VP, _VP]),
    "lfric_prec_destroy": (None, [_VP]),
    "lfric_linop_create_lma": (_I, [_VP, ctypes.POINTER(_VP)]),
    "lfric_linop_create_helmholtz": (_I, [_VP, _VP, _VP, _D, ctypes.POINTER(_VP)]),
    "lfric_linop_apply": (_I, [_VP, _VP, _VP]),
    "lfric_linop_destroy": (None, [_VP]),
    "lfric_prec_create_column_jacobi": (_I, [_VP, _I, ctypes.POINTER(_VP)]),
    "lfric_ugrid_read": (_I, [ctypes.c_char_p, ctypes.c_char_p, _VP, _VP, _VP, _VP]),
    "lfric_fs_create_ugrid": (_I, [ctypes.c_char_p, ctypes.c_char_p, _I, _I, _I, ctypes.POINTER(_VP)]),
    "lfric_cma_create": (_I, [_VP, _VP, _I, ctypes.POINTER(_VP)]),
    "lfric_cma_create_custom": (_I, [_VP, _VP, _I, _I, _I, _I, _I, ctypes.POINTER(_VP)]),
    "lfric_cma_info": (_I, [_VP, _VP]),
    "lfric_cma_upload": (_I, [_VP, _VP]),
    "lfric_cma_download": (_I, [_VP, _VP]),
    "lfric_cma_assemble": (_I, [_VP, _VP]),
    "lfric_cma_assemble_diag_hmht": (_I, [_VP, _VP, _VP]),
    "lfric_cma_assemble_m2v_lumped_inv": (_I, [_VP, _VP]),
    "lfric_cma_apply": (_I, [_VP, _VP, _VP, _I]),
    "lfric_cma_apply_inverse": (_I, [_VP, _VP, _VP]),
    "lfric_cma_multiply": (_I, [_VP, _VP, _VP]),
    "lfric_cma_scaled_add": (_I, [_VP, _VP, _VP, _D, _D]),
    "lfric_cma_destroy": (None, [_VP]),
    "lfric_prec_create_cma_inverse": (_I, [_VP, ctypes.POINTER(_VP)]),
    "lfric_tri_solve": (_I, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "lfric_gmres_solve": (_I, [_VP, _VP, _VP, _VP, ctypes.POINTER(SolverOpts), ctypes.POINTER(SolverRes)]),
    "lfric_ir_gmres_solve": (_I, [_VP, _VP, _VP, _VP, _VP, ctypes.POINTER(SolverOpts), ctypes.POINTER(SolverRes)]),
    "lfric_solver_opts_default": (None, [ctypes.POINTER(SolverOpts)]),
    "lfric_residual_norm": (_I, [_VP, _VP, _VP, _DP]),
    "lfric_cg_solve": (_I, [_VP, _VP, _VP, _VP, ctypes.POINTER(SolverOpts), ctypes.POINTER(SolverRes)]),
    "lfric_extrusion_eta": (_I, [_I, _I, _VP]),
    "lfric_set_stream": (_I, [_VP]),
    "lfric_lma_apply_stream": (_I, [_VP, _VP, _VP, _VP]),
    "lfric_fs_coordinates": (_I, [_VP, _I, _D, _D, _VP, _VP, _VP]),
    "lfric_fs_get_boundary_dofs": (_I, [_VP, _VP]),
    "lfric_fs_set_boundary_dofs": (_I, [_VP, _VP]),
    "lfric_dg_inc_apply": (_I, [_VP, _VP, _VP]),
    "lfric_enforce_bc": (_I, [_VP]),
    "lfric_linop_create_mass_matrix": (_I, [_VP, _I, ctypes.POINTER(_VP)]),
    "lfric_cg_solve_noalloc": (_I, [_VP, _VP, _VP, _VP, ctypes.POINTER(SolverOpts), ctypes.POINTER(SolverRes)]),
    "lfric_ir_cg_solve": (_I, [_VP, _VP, _VP, _VP, _VP, ctypes.POINTER(SolverOpts), ctypes.POINTER(SolverRes)]),
    "lfric_lma_apply_mrhs": (_I, [_VP, _VP, _VP, _I]),
    "lfric_op_tune": (_I, [_VP, _I, ctypes.POINTER(_I), _DP]),
    "lfric_op_set_window": (_I, [_VP, _I]),
    "lfric_op_time_apply": (_I, [_VP, _VP, _VP, _I, _DP]),
    "lfric_op_prepare_rows": (_I, [_VP]),
    "lfric_fgmres_solve": (_I, [_VP, _VP, _VP, _VP, ctypes.POINTER(SolverOpts), ctypes.POINTER(SolverRes)]),
    "lfric_bicgstab_solve": (_I, [_VP, _VP, _VP, _VP, ctypes.POINTER(SolverOpts), ctypes.POINTER(SolverRes)]),
    "lfric_gcr_solve": (_I, [_VP, _VP, _VP, _VP, ctypes.POINTER(SolverOpts), ctypes.POINTER(SolverRes)]),
    "lfric_jacobi_solve": (_I, [_VP, _VP, _VP, _VP, ctypes.POINTER(SolverOpts), ctypes.POINTER(SolverRes)]),
    "lfric_chebyshev_solve": (_I, [_VP, _VP, _VP, _VP, ctypes.POINTER(SolverOpts), ctypes.POINTER(SolverRes)]),
    "lfric_prec_only_solve": (_I, [_VP, _VP, _VP, _VP, ctypes.POINTER(SolverOpts), ctypes.POINTER(SolverRes)]),
    "lfric_solve": (_I, [_I, _VP, _VP, _VP, _VP, ctypes.POINTER(SolverOpts), ctypes.POINTER(SolverRes)]),
    "lfric_batch_create": (_I, [ctypes.POINTER(SolverOpts), ctypes.POINTER(_VP)]),
    "lfric_batch_add": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, ctypes.POINTER(_I)]),
    "lfric_batch_run": (_I, [_VP, _I, ctypes.POINTER(_I)]),
    "lfric_batch_solve": (_I, [_VP]),
    "lfric_batch_finished": (_I, [_VP, _I, ctypes.POINTER(_I)]),
    "lfric_batch_result": (_I, [_VP, _I, ctypes.POINTER(SolverRes)]),
    "lfric_batch_remove": (_I, [_VP, _I]),
    "lfric_batch_tune": (_I, [_VP, _DP]),
    "lfric_batch_config": (_I, [_VP, _I, _I, _I]),
    "lfric_batch_set_apply": (_I, [_VP, _I]),
    "lfric_batch_stats": (_I, [_VP, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_I), ctypes.POINTER(_I),
                               ctypes.POINTER(_I)]),
    "lfric_batch_destroy": (None, [_VP]),
}

_lib = None


def lib():
    """Load liblfric_hip.so; raise if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, allow=(LFRIC_OK,)):
    if rc in allow:
        return rc
    msg = lib().lfric_last_error().decode()
    if rc == LFRIC_NOT_CONVERGED:
        raise NotConverged(rc, msg)
    raise LfricError(rc, msg)


_initialised = None


def init(device=0):
    global _initialised
    if _initialised is None:
        check(lib().lfric_init(device))
        _initialised = device
    return _initialised


def synchronize():
    check(lib().lfric_synchronize())


def timer_start():
    check(lib().lfric_timer_start())


def timer_stop():
    ms = ctypes.c_double()
    check(lib().lfric_timer_stop(ctypes.byref(ms)))
    return ms.value


def device_bytes():
    return lib().lfric_device_bytes()


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def set_apply_form(form):
    """LMA apply used inside the solvers: 0 colour-ordered scatter, 1 gather (fused p.Ap in CG)."""
    check(lib().lfric_set_apply_form(int(form)))
