"""ctypes binding of libflgpu.so (the C-ABI declared in include/flgpu.h).

This is the same binding a femlisp maintainer writes with fl.port:def-function
(src/basic/port.lisp:401-455): opaque handles, int status, host arrays valid
for the call only.  Loading fails loudly if the library is not built; there is
no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FLGPU_LIB selects another build of the library (e.g. a variant for an A/B timing)
LIB_PATH = os.environ.get("FLGPU_LIB") or os.path.join(HERE, "libflgpu.so")

F64, F32 = 0, 1
ASM_F64, ASM_MFMA = 0, 1
SMOOTH_JACOBI, SMOOTH_SOR, SMOOTH_SSC, SMOOTH_PSC = 0, 1, 2, 3
OUTER_MG, OUTER_CG, OUTER_CG_PLAIN, OUTER_FCG = 0, 1, 2, 3
STOP_REDUCTION, STOP_DEFNORM = 0, 1
STORAGE_SELL, STORAGE_STENCIL, STORAGE_CELL = 0, 1, 2
TRANSFER_GEOMETRIC, TRANSFER_BLACKBOX = 0, 1
FALLBACK_NONE, FALLBACK_LU = 0, 1


class FlgpuError(RuntimeError):
    """Raised for a non-zero status, like femlisp's (error ...) on a failed
    foreign call (src/matlisp/compressed.lisp:327-333)."""

    def __init__(self, status, msg):
        super().__init__(f"flgpu status {status}: {msg}")
        self.status = status


class MgParams(ctypes.Structure):
    _fields_ = [("smoother", ctypes.c_int), ("omega", ctypes.c_double),
                ("pre_steps", ctypes.c_int), ("post_steps", ctypes.c_int),
                ("base_level", ctypes.c_int), ("symmetric", ctypes.c_int),
                ("prec", ctypes.c_int), ("galerkin", ctypes.c_int),
                ("transfer", ctypes.c_int), ("gamma", ctypes.c_int),
                ("additive", ctypes.c_int), ("fas", ctypes.c_int)]


class SolveParams(ctypes.Structure):
    _fields_ = [("outer", ctypes.c_int), ("alias", ctypes.c_int), ("stop", ctypes.c_int),
                ("tol", ctypes.c_double), ("maxsteps", ctypes.c_int),
                ("true_residual", ctypes.c_int), ("learned", ctypes.c_int),
                ("stall_window", ctypes.c_int), ("restart_cycle", ctypes.c_int),
                ("fallback", ctypes.c_int)]


class SolveReport(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("steps", ctypes.c_int),
                ("initial_defnorm", ctypes.c_double), ("final_defnorm", ctypes.c_double),
                ("true_residual_norm", ctypes.c_double), ("fell_back", ctypes.c_int),
                ("seconds", ctypes.c_double), ("fallback_used", ctypes.c_int)]


# name -> (restype, argtypes); every symbol of include/flgpu.h
_vp, _ci, _cl, _cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double
_P = ctypes.POINTER
SIGNATURES = {
    "flgpu_last_error": (_ci, [ctypes.c_char_p, _ci]),
    "flgpu_ctx_create": (_ci, [_ci, _P(_vp)]),
    "flgpu_ctx_destroy": (_ci, [_vp]),
    "flgpu_ctx_synchronize": (_ci, [_vp]),
    "flgpu_ctx_info": (_ci, [_vp, ctypes.c_char_p, _ci, _P(_ci), _P(_cl)]),
    "flgpu_ctx_timer_start": (_ci, [_vp]),
    "flgpu_ctx_timer_stop": (_ci, [_vp, _P(_cd)]),
    "flgpu_vec_create": (_ci, [_vp, _cl, _ci, _P(_vp)]),
    "flgpu_vec_destroy": (_ci, [_vp]),
    "flgpu_vec_info": (_ci, [_vp, _P(_cl), _P(_ci)]),
    "flgpu_vec_upload": (_ci, [_vp, _vp]),
    "flgpu_vec_download": (_ci, [_vp, _vp]),
    "flgpu_vec_fill": (_ci, [_vp, _cd]),
    "flgpu_vec_copy": (_ci, [_vp, _vp]),
    "flgpu_vec_scal": (_ci, [_cd, _vp]),
    "flgpu_vec_axpy": (_ci, [_cd, _vp, _vp]),
    "flgpu_vec_dot": (_ci, [_vp, _vp, _P(_cd)]),
    "flgpu_vec_dot_noalloc": (_ci, [_vp, _vp, _P(_cd)]),
    "flgpu_vec_norm2": (_ci, [_vp, _P(_cd)]),
    "flgpu_csr_upload": (_ci, [_vp, _ci, _ci, _cl, _vp, _vp, _vp, _ci, _P(_vp)]),
    "flgpu_csr_update_values": (_ci, [_vp, _vp]),
    "flgpu_csr_download": (_ci, [_vp, _vp, _vp, _vp]),
    "flgpu_mat_info": (_ci, [_vp, _P(_ci), _P(_ci), _P(_cl), _P(_ci)]),
    "flgpu_mat_convert": (_ci, [_vp, _ci, _P(_vp)]),
    "flgpu_mat_destroy": (_ci, [_vp]),
    "flgpu_mat_storage": (_ci, [_vp, _ci, _P(_ci)]),
    "flgpu_mat_stencil_planes": (_ci, [_vp, _P(_ci)]),
    "flgpu_gemm_nn": (_ci, [_cd, _vp, _vp, _cd, _vp]),
    "flgpu_gemm_tn": (_ci, [_cd, _vp, _vp, _cd, _vp]),
    "flgpu_compute_residual": (_ci, [_vp, _vp, _vp, _vp]),
    "flgpu_galerkin_product": (_ci, [_vp, _vp, _vp]),
    "flgpu_mat_set_colouring": (_ci, [_vp, _ci, _vp]),
    "flgpu_mat_colour_greedy": (_ci, [_vp, _P(_ci)]),
    "flgpu_jacobi_iterate": (_ci, [_vp, _cd, _vp, _vp]),
    "flgpu_jacobi_steps": (_ci, [_vp, _cd, _ci, _vp, _vp, _vp]),
    "flgpu_sor_iterate": (_ci, [_vp, _cd, _ci, _vp, _vp, _vp]),
    "flgpu_block_iterate": (_ci, [_vp, _ci, _cd, _ci, _vp, _vp, _vp]),
    "flgpu_lattice_assemble": (_ci, [_vp, _ci, _ci, _ci, _vp, _ci, _cd, _ci, _ci, _P(_vp), _P(_vp)]),
    "flgpu_lattice_reassemble": (_ci, [_vp, _vp, _ci, _ci]),
    "flgpu_lattice_interpolation": (_ci, [_vp, _ci, _ci, _ci, _ci, _P(_vp)]),
    "flgpu_lattice_projection": (_ci, [_vp, _ci, _ci, _ci, _ci, _P(_vp)]),
    "flgpu_lattice_colouring": (_ci, [_vp, _ci, _ci, _ci]),
    "flgpu_lattice_blackbox_interpolation": (_ci, [_vp, _vp]),
    "flgpu_lattice_element_matrices": (_ci, [_vp, _ci, _ci, _ci, _vp, _ci, _ci, _ci, _cl, _cl, _vp]),
    "flgpu_mg_create": (_ci, [_vp, _ci, _vp, _vp, _P(MgParams), _P(_vp)]),
    "flgpu_mg_create_lattice": (_ci, [_vp, _ci, _ci, _ci, _ci, _vp, _ci, _cd, _ci, _P(MgParams),
                                      _P(_vp), _P(_vp)]),
    "flgpu_mg_destroy": (_ci, [_vp]),
    "flgpu_mg_top_matrix": (_ci, [_vp, _P(_vp)]),
    "flgpu_mg_level_matrix": (_ci, [_vp, _ci, _ci, _P(_vp)]),
    "flgpu_mg_apply": (_ci, [_vp, _vp, _vp, _vp]),
    "flgpu_mg_use_graphs": (_ci, [_vp, _ci]),
    "flgpu_solve": (_ci, [_vp, _vp, _P(SolveParams), _vp, _vp, _P(SolveReport), _vp, _ci]),
    "flgpu_mg_reassemble": (_ci, [_vp, _vp, _ci]),
    "flgpu_unet_create": (_ci, [_vp, _ci, ctypes.c_ulong, _P(_vp)]),
    "flgpu_unet_destroy": (_ci, [_vp]),
    "flgpu_unet_nparams": (_ci, [_vp, _P(_cl)]),
    "flgpu_unet_get_params": (_ci, [_vp, _vp, _vp, _vp, _P(_ci)]),
    "flgpu_unet_set_params": (_ci, [_vp, _vp, _vp, _vp, _ci]),
    "flgpu_unet_forward": (_ci, [_vp, _ci, _ci, _vp, _vp]),
    "flgpu_unet_loss_grad": (_ci, [_vp, _ci, _ci, _vp, _vp, _P(_cd), _vp]),
    "flgpu_unet_train_step": (_ci, [_vp, _ci, _ci, _vp, _vp, _cd, _P(_cd)]),
    "flgpu_unet_pack": (_ci, [_vp, _vp, _ci, _vp, _cl, _ci]),
    "flgpu_vec_gather_rows": (_ci, [_vp, _cl, _vp, _ci, _vp]),
    "flgpu_mg_attach_unet": (_ci, [_vp, _vp, _vp, _ci]),
    "flgpu_system_read": (_ci, [_vp, ctypes.c_char_p, _ci, _P(_vp), _P(_vp)]),
    "flgpu_system_write": (_ci, [ctypes.c_char_p, _vp, _vp]),
}

_LIB = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run python -m gnu_femlisp_amd.build "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().flgpu_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(status):
    if status != 0:
        raise FlgpuError(status, last_error())
    return status


def ptr(a):
    """Pointer to a C-contiguous numpy array (None for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "host arrays must be C-contiguous"
    return a.ctypes.data
