"""ctypes binding of libcaffe_hip.so, the C ABI declared in include/caffe_hip.h.

The library is built in-tree by ``caffe_old_amd.build`` (hipcc, gfx950).  Loading
fails loudly when it is missing: there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcaffe_hip.so")

CAFFE_OK = 0
CAFFE_ERR_ARG = -1
CAFFE_ERR_HIP = -2
CAFFE_ERR_STATE = -3
CAFFE_ERR_NOT_CONVERGED = -4

SOLVE_PLAIN_FP64 = 0
SOLVE_IR_FP32 = 1
SOLVE_LEARNED_FCG = 2
SOLVE_MG = 3
SOLVE_SGS_FP64 = 4

# every symbol include/caffe_hip.h declares
EXPORTS = [
    "caffe_ctx_create", "caffe_ctx_destroy", "caffe_last_error", "caffe_case_setup", "caffe_set_solve_opts",
    "caffe_solve", "caffe_save_u0", "caffe_momentum_assemble", "caffe_momentum_solve", "caffe_rhiechow",
    "caffe_pcorr_assemble", "caffe_pcorr_solve", "caffe_correct", "caffe_ib_field", "caffe_ib_coeffs",
    "caffe_grad_u", "caffe_step_report", "caffe_last_reports", "caffe_last_courant", "caffe_field_download",
    "caffe_field_upload",
    "caffe_metric_download", "caffe_matrix_download", "caffe_rhs_download", "caffe_ghost_rows", "caffe_time_kernel",
    "caffe_kernel_probe", "caffe_kernel_probe_read",
    "caffe_sync",
    "caffe_padded_size", "caffe_device_bytes",
    # learned pressure-correction preconditioner
    "caffe_net_create", "caffe_net_destroy", "caffe_net_last_error", "caffe_net_param_count", "caffe_net_get_params",
    "caffe_net_set_params", "caffe_net_forward", "caffe_net_backward", "caffe_net_precond", "caffe_net_loss_grad",
    "caffe_net_set_post_sweeps",
    "caffe_net_collect", "caffe_net_clear", "caffe_net_dataset_size", "caffe_net_train", "caffe_net_eval",
    "caffe_ctx_attach_net", "caffe_ctx_set_fcg_coarse", "caffe_ctx_set_fcg_post_sweeps", "caffe_ctx_set_fcg_guess",
    # SparseMatrix / SparseVector compatibility path
    "caffe_matrix_create", "caffe_matrix_destroy", "caffe_matrix_last_error", "caffe_matrix_set_rows",
    "caffe_matrix_set_value", "caffe_matrix_zero_entries", "caffe_matrix_assemble", "caffe_matrix_solve",
    "caffe_matrix_solve_noalloc",
    "caffe_matrix_last_report", "caffe_matrix_multiply", "caffe_matrix_multiply_add", "caffe_matrix_scale",
    "caffe_vector_create", "caffe_vector_destroy", "caffe_vector_last_error", "caffe_vector_set_values",
    "caffe_vector_get_values", "caffe_vector_zero_entries", "caffe_vector_scale", "caffe_vector_norm",
]


class CaffeCase(C.Structure):
    _fields_ = [
        ("nI", C.c_int), ("nJ", C.c_int), ("nK", C.c_int),
        ("algo", C.c_int), ("steady", C.c_int),
        ("rho", C.c_double), ("mu", C.c_double), ("alphaU", C.c_double), ("alphaP", C.c_double),
        ("nInner", C.c_int), ("nPCorr", C.c_int),
        ("bcType", C.c_int * 6), ("bcVec", C.c_double * 18), ("bcVal", C.c_double * 6),
        ("uInit", C.c_double * 3), ("pInit", C.c_double),
        ("sphereR", C.c_double), ("sphereC", C.c_double * 3),
        ("uniform", C.c_int), ("nonOrth", C.c_int),
    ]


class SolveOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("rtol", C.c_double), ("maxit", C.c_int), ("inner_rtol", C.c_double)]


class SolveReport(C.Structure):
    _fields_ = [("iters", C.c_int), ("refinements", C.c_int), ("relres", C.c_double), ("converged", C.c_int),
                ("fell_back", C.c_int), ("ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class CaffeError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise CaffeError(f"{path} is not built: run __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(path)
    vp, dp, ip, lp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
    fp = C.POINTER(C.c_float)
    sig = {
        "caffe_ctx_create": ([C.c_int, C.POINTER(C.c_void_p)], C.c_int),
        "caffe_ctx_destroy": ([vp], C.c_int),
        "caffe_last_error": ([vp], C.c_char_p),
        "caffe_case_setup": ([vp, C.POINTER(CaffeCase), dp], C.c_int),
        "caffe_set_solve_opts": ([vp, C.c_int, C.POINTER(SolveOpts)], C.c_int),
        "caffe_solve": ([vp, C.c_double, dp], C.c_int),
        "caffe_save_u0": ([vp], C.c_int),
        "caffe_momentum_assemble": ([vp, C.c_double], C.c_int),
        "caffe_momentum_solve": ([vp, C.POINTER(SolveReport)], C.c_int),
        "caffe_rhiechow": ([vp], C.c_int),
        "caffe_pcorr_assemble": ([vp], C.c_int),
        "caffe_pcorr_solve": ([vp, C.POINTER(SolveOpts), C.POINTER(SolveReport)], C.c_int),
        "caffe_correct": ([vp], C.c_int),
        "caffe_ib_field": ([vp], C.c_int),
        "caffe_ib_coeffs": ([vp], C.c_int),
        "caffe_grad_u": ([vp], C.c_int),
        "caffe_step_report": ([vp, C.c_double, dp, dp, dp], C.c_int),
        "caffe_last_reports": ([vp, C.POINTER(SolveReport), C.POINTER(SolveReport)], C.c_int),
        "caffe_last_courant": ([vp, dp], C.c_int),
        "caffe_field_download": ([vp, C.c_char_p, dp, C.c_long], C.c_int),
        "caffe_field_upload": ([vp, C.c_char_p, dp, C.c_long], C.c_int),
        "caffe_metric_download": ([vp, C.c_char_p, dp, C.c_long], C.c_int),
        "caffe_matrix_download": ([vp, C.c_int, dp, C.c_long], C.c_int),
        "caffe_rhs_download": ([vp, C.c_int, dp, C.c_long], C.c_int),
        "caffe_ghost_rows": ([vp, ip, lp, lp, dp, dp], C.c_int),
        "caffe_time_kernel": ([vp, C.c_char_p, C.c_int, dp, dp], C.c_int),
        "caffe_kernel_probe": ([vp, C.c_char_p, C.c_int], C.c_int),
        "caffe_kernel_probe_read": ([vp, dp, lp, dp], C.c_int),
        "caffe_sync": ([vp], C.c_int),
        "caffe_padded_size": ([vp], C.c_long),
        "caffe_device_bytes": ([vp], C.c_long),
        "caffe_net_create": ([C.c_int, C.c_ulonglong, C.POINTER(C.c_void_p)], C.c_int),
        "caffe_net_destroy": ([vp], C.c_int),
        "caffe_net_last_error": ([vp], C.c_char_p),
        "caffe_net_param_count": ([], C.c_long),
        "caffe_net_get_params": ([vp, fp, C.c_long], C.c_int),
        "caffe_net_set_params": ([vp, fp, C.c_long], C.c_int),
        "caffe_net_forward": ([vp, C.c_int, C.c_int, C.c_int, fp, fp], C.c_int),
        "caffe_net_backward": ([vp, C.c_int, C.c_int, C.c_int, fp, fp, fp], C.c_int),
        "caffe_net_precond": ([vp, C.c_int, C.c_int, C.c_int, fp, fp, fp], C.c_int),
        "caffe_net_set_post_sweeps": ([vp, C.c_int], C.c_int),
        "caffe_net_loss_grad": ([vp, C.c_int, C.c_int, C.c_int, fp, fp, fp, dp, fp], C.c_int),
        "caffe_net_collect": ([vp, vp, lp], C.c_int),
        "caffe_net_clear": ([vp], C.c_int),
        "caffe_net_dataset_size": ([vp], C.c_long),
        "caffe_net_train": ([vp, C.c_int, C.c_int, C.c_float, C.c_double, C.c_ulonglong, fp], C.c_int),
        "caffe_net_eval": ([vp, fp, C.c_long], C.c_int),
        "caffe_ctx_attach_net": ([vp, vp], C.c_int),
        "caffe_ctx_set_fcg_coarse": ([vp, C.c_int], C.c_int),
        "caffe_ctx_set_fcg_post_sweeps": ([vp, C.c_int], C.c_int),
        "caffe_ctx_set_fcg_guess": ([vp, C.c_int], C.c_int),
        "caffe_matrix_create": ([C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                 C.POINTER(C.c_void_p)], C.c_int),
        "caffe_matrix_destroy": ([vp], C.c_int),
        "caffe_matrix_last_error": ([vp], C.c_char_p),
        "caffe_matrix_set_rows": ([vp, C.c_int, ip, ip, ip, dp], C.c_int),
        "caffe_matrix_set_value": ([vp, C.c_int, C.c_int, C.c_double], C.c_int),
        "caffe_matrix_zero_entries": ([vp], C.c_int),
        "caffe_matrix_assemble": ([vp], C.c_int),
        "caffe_matrix_solve": ([vp, vp, vp, ip], C.c_int),
        "caffe_matrix_solve_noalloc": ([vp, vp, vp, ip], C.c_int),
        "caffe_matrix_last_report": ([vp, C.POINTER(SolveReport)], C.c_int),
        "caffe_matrix_multiply": ([vp, vp, vp], C.c_int),
        "caffe_matrix_multiply_add": ([vp, vp, vp, vp], C.c_int),
        "caffe_matrix_scale": ([C.c_double, vp], C.c_int),
        "caffe_vector_create": ([C.c_int, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
        "caffe_vector_destroy": ([vp], C.c_int),
        "caffe_vector_last_error": ([vp], C.c_char_p),
        "caffe_vector_set_values": ([vp, C.c_int, ip, dp], C.c_int),
        "caffe_vector_get_values": ([vp, C.c_int, ip, dp], C.c_int),
        "caffe_vector_zero_entries": ([vp], C.c_int),
        "caffe_vector_scale": ([C.c_double, vp], C.c_int),
        "caffe_vector_norm": ([vp, C.c_int, dp], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L
