"""ctypes binding of libsuperman_amd.so (the C ABI in include/superman_amd.h).

The library is loaded from the package directory only (built in-tree by
superman_amd.build).  There is no fallback: if the library is missing or a
symbol is absent, loading raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsuperman_amd.so")

SPM_OK = 0
SPM_EINVAL = -1
SPM_EHIP = -2
SPM_ENOMEM = -3
SPM_ERANGE = -4
SPM_ECOMPILE = -5
SPM_ENOTINT = -6
SPM_ECANCELLED = -7
SPM_EAGAIN = -8
SPM_ETIMEDOUT = -9
SPM_EFAULTINJ = -10

SPM_CG_REGISTER_EFFICIENT = 1

SPM_PREC_PLAIN = 0
SPM_PREC_COMPENSATED = 1
SPM_PREC_F32_PRODUCT = 2

SPM_PATH_INT = 1
SPM_PATH_GEN_INT = 2
SPM_PATH_SCALED_INT = 3
SPM_PATH_GEN_REAL = 4
SPM_PATH_DENSE = 5

SPM_INT_WALK_AUTO = 0
SPM_INT_WALK_INT32 = 1
SPM_INT_WALK_FP64 = 2
SPM_INT_WALK_SPARSE = 3

SPM_DECOMP_DM = 1
SPM_DECOMP_COMPONENTS = 2
SPM_DECOMP_COMPRESS = 4
SPM_DECOMP_CODEGEN = 8
SPM_DECOMP_SMALL = 16
SPM_DECOMP_DEFAULT = 7

SPM_KIND_REAL = 1
SPM_KIND_COMPLEX = 2
SPM_KIND_INT = 3
SPM_BATCH_MAX_N = 28
SPM_EXEC_DEVICE = 0
SPM_EXEC_DRYRUN = 1


class SpmResult(ctypes.Structure):
    _fields_ = [
        ("re", ctypes.c_double), ("im", ctypes.c_double),
        ("hi_re", ctypes.c_double), ("lo_re", ctypes.c_double),
        ("hi_im", ctypes.c_double), ("lo_im", ctypes.c_double),
        ("abs_sum", ctypes.c_double), ("err_est", ctypes.c_double),
        ("exact_lo", ctypes.c_int64), ("exact_hi", ctypes.c_int64),
        ("is_exact", ctypes.c_int32), ("status", ctypes.c_int32),
        ("seconds", ctypes.c_double), ("kernel_ms", ctypes.c_double),
        ("subsets", ctypes.c_uint64),
        ("n", ctypes.c_int32), ("launches", ctypes.c_int32),
    ]


class SpmBatchConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int), ("slots", ctypes.c_int), ("max_requests", ctypes.c_int),
        ("max_items", ctypes.c_int), ("arena_bytes", ctypes.c_size_t), ("max_pending", ctypes.c_int),
        ("max_batch_work", ctypes.c_double), ("use_graph", ctypes.c_int), ("executor", ctypes.c_int),
    ]


class SpmBatchTuning(ctypes.Structure):
    _fields_ = [("fused", ctypes.c_int), ("waves_per_cu", ctypes.c_int), ("log2_chunk", ctypes.c_int * 4)]


class SpmRequest(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_int), ("a", ctypes.POINTER(ctypes.c_double)),
                ("tag", ctypes.c_uint64), ("deadline_s", ctypes.c_double)]


class SpmCompletion(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("result", SpmResult),
                ("t_submit", ctypes.c_double), ("t_launch", ctypes.c_double), ("t_done", ctypes.c_double)]


class SpmServeStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("wall_s", ctypes.c_double), ("perms_per_s", ctypes.c_double), ("p50_s", ctypes.c_double),
                ("p99_s", ctypes.c_double), ("mean_s", ctypes.c_double), ("max_s", ctypes.c_double),
                ("mean_batch", ctypes.c_double), ("device_busy", ctypes.c_double), ("late_publish", ctypes.c_uint64),
                ("retried", ctypes.c_uint64)]


class SpmError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = lib().spm_strerror(status).decode() if _lib is not None else str(status)
        super().__init__(f"{what}: {msg} ({status})" if what else f"{msg} ({status})")


_lib = None
_lock = threading.Lock()

_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I32 = ctypes.POINTER(ctypes.c_int32)
_VP = ctypes.c_void_p
_R = ctypes.POINTER(SpmResult)

# name -> (restype, argtypes); mirrors include/superman_amd.h
PROTOTYPES = {
    "spm_version": (ctypes.c_char_p, []),
    "spm_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "spm_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "spm_device_synchronize": (ctypes.c_int, [ctypes.c_int]),
    "spm_device_info": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.c_char_p, ctypes.c_int]),
    "spm_dense_f64": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _R]),
    "spm_dense_f64_noalloc": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _R]),
    "spm_dense_c128": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _R]),
    "spm_int_exact": (ctypes.c_int, [_I64, ctypes.c_int, ctypes.c_int, _VP, _R]),
    "spm_plan_int": (ctypes.c_int, [_I64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_VP)]),
    "spm_plan_int_walk": (ctypes.c_int, [_I64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_VP)]),
    "spm_sparse_codegen": (ctypes.c_int, [_I64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _R]),
    "spm_codegen_source": (ctypes.c_int, [_I64, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t)]),
    "spm_codegen_compile_only": (ctypes.c_int, [_I64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                                ctypes.POINTER(ctypes.c_size_t)]),
    "spm_sparse_codegen_real": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _R]),
    "spm_codegen_source_real": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                               ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "spm_plan_codegen_real": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(_VP)]),
    "spm_permanent_auto": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _R,
                                          ctypes.POINTER(ctypes.c_int)]),
    "spm_choose_path": (ctypes.c_int, [_D, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "spm_path_estimate": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "spm_ut_order": (ctypes.c_int, [_D, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "spm_plan_codegen": (ctypes.c_int, [_I64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_VP)]),
    "spm_plan_codegen_info": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_uint64)]),
    "spm_plan_dense_f64": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_VP)]),
    "spm_plan_dense_c128": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_VP)]),
    "spm_plan_run": (ctypes.c_int, [_VP, _VP, _R]),
    "spm_plan_destroy": (ctypes.c_int, [_VP]),
    "spm_plan_info": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64)]),
    "spm_plan_kernel_name": (ctypes.c_char_p, [_VP]),
    "spm_plan_set_log2_chunk": (ctypes.c_int, [_VP, ctypes.c_int]),
    "spm_gray_geometry": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_int)]),
    "spm_decompose": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_VP)]),
    "spm_decomp_stats": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_int)]),
    "spm_decomp_leaf": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _D]),
    "spm_decomp_destroy": (ctypes.c_int, [_VP]),
    "spm_permanent_decomposed": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _R]),
    "spm_permanent_decomposed_scaled": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                       ctypes.c_double, ctypes.c_int, _VP, _R]),
    "spm_sinkhorn_scale": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_double, _D, _D]),
    "spm_batch_config_default": (None, [ctypes.POINTER(SpmBatchConfig)]),
    "spm_batch_create": (ctypes.c_int, [ctypes.POINTER(SpmBatchConfig), ctypes.POINTER(_VP)]),
    "spm_batch_destroy": (ctypes.c_int, [_VP]),
    "spm_batch_get_tuning": (ctypes.c_int, [_VP, ctypes.POINTER(SpmBatchTuning)]),
    "spm_batch_set_tuning": (ctypes.c_int, [_VP, ctypes.POINTER(SpmBatchTuning)]),
    "spm_batch_autotune": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(SpmRequest), ctypes.c_int,
                                          ctypes.POINTER(SpmBatchTuning), ctypes.POINTER(ctypes.c_double)]),
    "spm_batch_run": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(SpmRequest), _R]),
    "spm_serve_submit": (ctypes.c_int, [_VP, ctypes.POINTER(SpmRequest), ctypes.POINTER(ctypes.c_uint64)]),
    "spm_serve_progress": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int)]),
    "spm_serve_poll": (ctypes.c_int, [_VP, ctypes.POINTER(SpmCompletion), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "spm_serve_cancel": (ctypes.c_int, [_VP, ctypes.c_uint64]),
    "spm_serve_pending": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "spm_serve_replay": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(SpmRequest), _D, _R, _D,
                                        ctypes.POINTER(SpmServeStats)]),
    "spm_batch_counters": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "spm_batch_inject_fault": (ctypes.c_int, [_VP, ctypes.c_int]),
    "spm_batch_inject_bad_results": (ctypes.c_int, [_VP, ctypes.c_int]),
    "spm_batch_last_error": (ctypes.c_char_p, [_VP]),
}


def lib() -> ctypes.CDLL:
    """Load the in-tree library (raises OSError if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise OSError(f"{LIB_PATH} not found: build it with `python -m superman_amd.build`")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in PROTOTYPES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(status: int, what: str = "") -> None:
    if status != SPM_OK:
        raise SpmError(status, what)
