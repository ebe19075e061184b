"""ctypes loader for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference answer in tests and as the CPU
baseline.  The product package (superman_amd) never imports this module.

Every function takes a numpy array and returns Python numbers; see
ryser_oracle.c for the reference file:line each C function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build(quiet: bool = True) -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    out = subprocess.run(["make", "-C", HERE, "liboracle.so"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int64)
        L.orc_dense_ref.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        L.orc_dense_ref.restype = ctypes.c_double
        L.orc_dense_ref_sample.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_longlong]
        L.orc_dense_ref_sample.restype = ctypes.c_double
        L.orc_sparse_ref.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        L.orc_sparse_ref.restype = ctypes.c_double
        L.orc_complex_ref.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp]
        L.orc_complex_ref.restype = None
        L.orc_complex_ref_sample.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_longlong, dp]
        L.orc_complex_ref_sample.restype = None
        L.orc_dense_dd.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_dense_dd.restype = None
        L.orc_complex_dd.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp]
        L.orc_complex_dd.restype = None
        L.orc_int_exact.argtypes = [ip, ctypes.c_int, ctypes.c_int, ip]
        L.orc_int_exact.restype = ctypes.c_int
        L.orc_int_log2_bound.argtypes = [ip, ctypes.c_int]
        L.orc_int_log2_bound.restype = ctypes.c_double
        L.orc_ascending_col_sort.argtypes = [dp, ctypes.c_int, dp]
        L.orc_ascending_col_sort.restype = None
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _real(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError("square matrix expected")
    return a


def _cplx(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError("square matrix expected")
    return a


def dense_ref(a, threads: int = 1) -> float:
    """Restated dpNaivePerman (reference CPU dense path)."""
    a = _real(a)
    return lib().orc_dense_ref(_dptr(a), a.shape[0], threads)


def dense_ref_sample(a, threads: int, count: int) -> float:
    """The restated reference walk over Gray indices [1, 1+count) only
    (bounded CPU-baseline sample); returns the unscaled partial sum."""
    a = _real(a)
    return lib().orc_dense_ref_sample(_dptr(a), a.shape[0], threads, count)


def sparse_ref(a, threads: int = 1) -> float:
    """Restated ascendingColSort + spNaivePerman (reference CPU sparse path)."""
    a = _real(a)
    return lib().orc_sparse_ref(_dptr(a), a.shape[0], threads)


def ascending_col_sort(a) -> np.ndarray:
    a = _real(a)
    out = np.empty_like(a)
    lib().orc_ascending_col_sort(_dptr(a), a.shape[0], _dptr(out))
    return out


def complex_ref(a, threads: int = 1) -> complex:
    """Restated cpuComputeComplex (reference CPU complex path)."""
    a = _cplx(a)
    out = np.zeros(2)
    lib().orc_complex_ref(_dptr(a.view(np.float64)), a.shape[0], threads, _dptr(out))
    return complex(out[0], out[1])


def complex_ref_sample(a, threads: int, count: int) -> complex:
    """The restated reference complex walk over Gray indices [1, 1+count)
    only (bounded CPU-baseline sample); returns the unscaled partial sum."""
    a = _cplx(a)
    out = np.zeros(2)
    lib().orc_complex_ref_sample(_dptr(a.view(np.float64)), a.shape[0], threads, count, _dptr(out))
    return complex(out[0], out[1])


def dense_dd(a, threads: int = 8) -> tuple[float, float]:
    """Double-double Ryser: returns (hi, lo) with perm ~= hi + lo."""
    a = _real(a)
    hi = np.zeros(1)
    lo = np.zeros(1)
    lib().orc_dense_dd(_dptr(a), a.shape[0], threads, _dptr(hi), _dptr(lo))
    return float(hi[0]), float(lo[0])


def complex_dd(a, threads: int = 8) -> tuple[float, float, float, float]:
    """Double-double complex Ryser: (re_hi, re_lo, im_hi, im_lo)."""
    a = _cplx(a)
    out = np.zeros(4)
    lib().orc_complex_dd(_dptr(a.view(np.float64)), a.shape[0], threads, _dptr(out))
    return tuple(float(v) for v in out)


def int_exact(a, threads: int = 8) -> tuple[int, int]:
    """Exact integer Ryser.  Returns (perm, status); status 0 = exact,
    1 = only known mod 2^128 (bound too large), negative = error."""
    a = np.ascontiguousarray(a, dtype=np.int64)
    out = np.zeros(2, dtype=np.int64)
    st = lib().orc_int_exact(a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), a.shape[0], threads,
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    lo = int(out[0]) & ((1 << 64) - 1)
    hi = int(out[1])
    return (hi << 64) | lo, st


def int_log2_bound(a) -> float:
    a = np.ascontiguousarray(a, dtype=np.int64)
    return lib().orc_int_log2_bound(a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), a.shape[0])


def sinkhorn_scale(a, iterations: int = 100, target: float = 2.0):
    """Restates IO::scale (IO/IO.h:109-161) in plain Python floats (small
    matrices only): `iterations` rounds of row then column balancing, each
    line scale = target / sum_k (a * r_i) * c_j where that sum is nonzero,
    then the final in-place pass over (i, j) in row order multiplying a[i][j]
    by r_i and a[j][i] by c_i.  Returns (scaled, r, c)."""
    a = [[float(v) for v in row] for row in np.asarray(a, dtype=np.float64)]
    n = len(a)
    r, c = [1.0] * n, [1.0] * n
    for _ in range(iterations):
        for i in range(n):
            s = 0.0
            for j in range(n):
                s += a[i][j] * r[i] * c[j]
            if s != 0:
                r[i] = float(target) / s
        for j in range(n):
            s = 0.0
            for i in range(n):
                s += a[i][j] * r[i] * c[j]
            if s != 0:
                c[j] = float(target) / s
    for i in range(n):
        for j in range(n):
            a[i][j] *= r[i]
            a[j][i] *= c[i]
    return np.array(a), np.array(r), np.array(c)
