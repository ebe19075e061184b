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
    "spm_dense_c128": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _R]),
    "spm_int_exact": (ctypes.c_int, [_I64, ctypes.c_int, ctypes.c_int, _VP, _R]),
    "spm_plan_int": (ctypes.c_int, [_I64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_VP)]),
    "spm_sparse_codegen": (ctypes.c_int, [_I64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _R]),
    "spm_codegen_source": (ctypes.c_int, [_I64, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t)]),
    "spm_codegen_compile_only": (ctypes.c_int, [_I64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                                ctypes.POINTER(ctypes.c_size_t)]),
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
    "spm_gray_geometry": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_int)]),
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
