"""Python host API over the C ABI.

    permanent(a, precision="kahan", device=0) -> Permanent

dispatches like the reference's AlgorithmSelector for the GPU modes
(CommonFiles/AlgorithmSelector.h:48-141): real matrices go to the dense
Gray-code Ryser kernel.  Precision names follow the reference's
`calculation_precision` values (IO/IO.h:358-380): "dd" = plain FP64,
"kahan" = compensated (the reference default).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import SpmResult, check, lib

PRECISIONS = {"dd": _lib.SPM_PREC_PLAIN, "plain": _lib.SPM_PREC_PLAIN,
              "kahan": _lib.SPM_PREC_COMPENSATED, "compensated": _lib.SPM_PREC_COMPENSATED,
              "f32_product": _lib.SPM_PREC_F32_PRODUCT}


@dataclass
class Permanent:
    value: complex | float
    hi: complex | float
    lo: complex | float
    abs_sum: float
    err_est: float
    exact: int | None
    seconds: float
    kernel_ms: float
    subsets: int
    n: int
    launches: int

    @property
    def cond(self) -> float:
        v = abs(self.value)
        return self.abs_sum / v if v else float("inf")


def _to_perm(r: SpmResult, cplx: bool) -> Permanent:
    exact = None
    if r.is_exact:
        exact = (int(r.exact_hi) << 64) | (int(r.exact_lo) & ((1 << 64) - 1))
    if cplx:
        value, hi, lo = complex(r.re, r.im), complex(r.hi_re, r.hi_im), complex(r.lo_re, r.lo_im)
    else:
        value, hi, lo = r.re, r.hi_re, r.lo_re
    return Permanent(value, hi, lo, r.abs_sum, r.err_est, exact, r.seconds, r.kernel_ms,
                     int(r.subsets), int(r.n), int(r.launches))


def _prec(p) -> int:
    if isinstance(p, int):
        return p
    try:
        return PRECISIONS[p]
    except KeyError:
        raise ValueError(f"unknown precision {p!r}") from None


def _square(a, dtype) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError("square matrix expected")
    return a


def dense_f64(a, precision="kahan", device: int = 0, stream: int | None = None) -> Permanent:
    a = _square(a, np.float64)
    r = SpmResult()
    st = lib().spm_dense_f64(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0],
                             _prec(precision), device, stream, ctypes.byref(r))
    check(st, "spm_dense_f64")
    return _to_perm(r, False)


def dense_f64_noalloc(a, precision="kahan", device: int = 0, stream: int | None = None) -> Permanent:
    """dense_f64 through spm_dense_f64_noalloc: the same result with no
    device allocation per call (one fixed per-device array; calls that
    overlap on one device race)."""
    a = _square(a, np.float64)
    r = SpmResult()
    st = lib().spm_dense_f64_noalloc(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0],
                                     _prec(precision), device, stream, ctypes.byref(r))
    check(st, "spm_dense_f64_noalloc")
    return _to_perm(r, False)


def dense_c128(a, precision="kahan", device: int = 0, stream: int | None = None) -> Permanent:
    a = _square(a, np.complex128)
    r = SpmResult()
    st = lib().spm_dense_c128(a.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0],
                              _prec(precision), device, stream, ctypes.byref(r))
    check(st, "spm_dense_c128")
    return _to_perm(r, True)


def int_exact(a, device: int = 0, stream: int | None = None) -> Permanent:
    """Exact permanent of an integer matrix (result in .exact)."""
    a = np.asarray(a)
    if a.dtype.kind == "f":
        if not np.all(np.isfinite(a)) or not np.all(a == np.round(a)):
            raise ValueError("int_exact needs integer entries")
    a = _square(a, np.int64)
    r = SpmResult()
    st = lib().spm_int_exact(a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), a.shape[0], device, stream,
                             ctypes.byref(r))
    check(st, "spm_int_exact")
    return _to_perm(r, False)


def is_integral(a) -> bool:
    a = np.asarray(a)
    if np.iscomplexobj(a):
        return False
    if a.dtype.kind in "iub":
        return True
    return bool(np.all(np.isfinite(a)) and np.all(a == np.round(a)) and np.all(np.abs(a) <= 2 ** 30))


def permanent(a, precision="auto", device: int = 0) -> Permanent:
    """Permanent of a square matrix on the GPU.

    precision: "auto" (integer matrices exactly when their bound allows, else
    compensated FP64 / complex128), "exact", "kahan" (compensated) or "dd"
    (plain FP64, the reference's name)."""
    a = np.asarray(a)
    if np.iscomplexobj(a):
        return dense_c128(a, "kahan" if precision in ("auto", "exact") else precision, device)
    if precision == "exact":
        return int_exact(a, device)
    if precision == "auto":
        return permanent_auto(a, "kahan", device)[0]
    return dense_f64(a, precision, device)


PATHS = {_lib.SPM_PATH_INT: "int", _lib.SPM_PATH_GEN_INT: "gen_int", _lib.SPM_PATH_SCALED_INT: "scaled_int",
         _lib.SPM_PATH_GEN_REAL: "gen_real", _lib.SPM_PATH_DENSE: "dense"}


def permanent_auto(a, precision="kahan", device: int = 0, stream: int | None = None) -> tuple[Permanent, str]:
    """The automatically chosen path (spm_permanent_auto): (result, path name)."""
    a = _square(a, np.float64)
    r = SpmResult()
    path = ctypes.c_int()
    check(lib().spm_permanent_auto(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0], _prec(precision),
                                   device, stream, ctypes.byref(r), ctypes.byref(path)), "spm_permanent_auto")
    return _to_perm(r, False), PATHS[path.value]


def choose_path(a) -> str:
    """The path spm_permanent_auto would take (host only, no device)."""
    a = _square(a, np.float64)
    path = ctypes.c_int()
    check(lib().spm_choose_path(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0], ctypes.byref(path)),
          "spm_choose_path")
    return PATHS[path.value]


def path_estimate(a, path: str) -> float:
    """Estimated seconds of one path for a (compile time included)."""
    a = _square(a, np.float64)
    inv = {v: k for k, v in PATHS.items()}
    sec = ctypes.c_double()
    check(lib().spm_path_estimate(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0], inv[path],
                                  ctypes.byref(sec)), "spm_path_estimate")
    return sec.value


class DensePlan:
    """A permanent with the matrix resident on the device: FP64, complex128
    for a complex matrix, or exact integer with precision="exact"."""

    INT_WALKS = {"auto": _lib.SPM_INT_WALK_AUTO, "int32": _lib.SPM_INT_WALK_INT32, "fp64": _lib.SPM_INT_WALK_FP64,
                 "sparse": _lib.SPM_INT_WALK_SPARSE}

    def __init__(self, a, precision="kahan", device: int = 0, int_walk: str = "auto"):
        self.complex = np.iscomplexobj(a)
        self._h = ctypes.c_void_p()
        if precision == "exact":
            a = _square(a, np.int64)
            check(lib().spm_plan_int_walk(a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), a.shape[0],
                                          self.INT_WALKS[int_walk], device, ctypes.byref(self._h)),
                  "spm_plan_int_walk")
        elif self.complex:
            a = _square(a, np.complex128)
            check(lib().spm_plan_dense_c128(a.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            a.shape[0], _prec(precision), device, ctypes.byref(self._h)),
                  "spm_plan_dense_c128")
        else:
            a = _square(a, np.float64)
            check(lib().spm_plan_dense_f64(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0],
                                           _prec(precision), device, ctypes.byref(self._h)), "spm_plan_dense_f64")
        self.n = a.shape[0]

    def run(self, stream: int | None = None) -> Permanent:
        r = SpmResult()
        check(lib().spm_plan_run(self._h, stream, ctypes.byref(r)), "spm_plan_run")
        return _to_perm(r, self.complex)

    def set_log2_chunk(self, log2_chunk: int) -> None:
        """Dense plans: walk chunks of 2^log2_chunk Gray steps (same
        permanent, other rounding)."""
        check(lib().spm_plan_set_log2_chunk(self._h, int(log2_chunk)), "spm_plan_set_log2_chunk")

    def info(self) -> dict:
        g, b, l = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        t = ctypes.c_uint64()
        check(lib().spm_plan_info(self._h, ctypes.byref(g), ctypes.byref(b), ctypes.byref(l), ctypes.byref(t)))
        return {"grid": g.value, "block": b.value, "log2_chunk": l.value, "tasks": t.value,
                "kernel": lib().spm_plan_kernel_name(self._h).decode()}

    def close(self):
        if self._h:
            lib().spm_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def device_count() -> int:
    c = ctypes.c_int()
    check(lib().spm_device_count(ctypes.byref(c)))
    return c.value


def synchronize(device: int = 0) -> None:
    check(lib().spm_device_synchronize(device), "spm_device_synchronize")


def device_info(device: int = 0) -> dict:
    cu, clk = ctypes.c_int(), ctypes.c_int()
    name = ctypes.create_string_buffer(256)
    check(lib().spm_device_info(device, ctypes.byref(cu), ctypes.byref(clk), name, 256), "spm_device_info")
    return {"compute_units": cu.value, "clock_khz": clk.value, "name": name.value.decode()}


# ---------------------------------------------------------------- SpaRyser codegen
def _int_square(a) -> np.ndarray:
    a = np.asarray(a)
    if a.dtype.kind == "f" and not np.all(a == np.round(a)):
        raise ValueError("generated SpaRyser kernels need integer entries")
    return _square(a, np.int64)


def sparse_codegen(a, register_efficient: bool = True, device: int = 0, stream: int | None = None) -> Permanent:
    """Exact permanent through a kernel generated for a's pattern (hiprtc)."""
    a = _int_square(a)
    r = SpmResult()
    opt = _lib.SPM_CG_REGISTER_EFFICIENT if register_efficient else 0
    check(lib().spm_sparse_codegen(a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), a.shape[0], opt, device, stream,
                                   ctypes.byref(r)), "spm_sparse_codegen")
    return _to_perm(r, False)


def codegen_source(a, register_efficient: bool = True) -> str:
    a = _int_square(a)
    opt = _lib.SPM_CG_REGISTER_EFFICIENT if register_efficient else 0
    need = ctypes.c_size_t()
    p = a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    check(lib().spm_codegen_source(p, a.shape[0], opt, None, 0, ctypes.byref(need)), "spm_codegen_source")
    buf = ctypes.create_string_buffer(need.value)
    check(lib().spm_codegen_source(p, a.shape[0], opt, buf, need.value, None), "spm_codegen_source")
    return buf.value.decode()


def codegen_compile_only(a, register_efficient: bool = True) -> dict:
    a = _int_square(a)
    opt = _lib.SPM_CG_REGISTER_EFFICIENT if register_efficient else 0
    sec, nbytes = ctypes.c_double(), ctypes.c_size_t()
    check(lib().spm_codegen_compile_only(a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), a.shape[0], opt,
                                         ctypes.byref(sec), ctypes.byref(nbytes)), "spm_codegen_compile_only")
    return {"seconds": sec.value, "code_bytes": nbytes.value}


def sparse_codegen_real(a, precision="kahan", register_efficient: bool = True, device: int = 0,
                        stream: int | None = None) -> Permanent:
    """FP64 permanent through a kernel generated for a's pattern and exact
    values (hiprtc): flips update only the flipped column's rows."""
    a = _square(a, np.float64)
    r = SpmResult()
    opt = _lib.SPM_CG_REGISTER_EFFICIENT if register_efficient else 0
    check(lib().spm_sparse_codegen_real(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0], opt,
                                        _prec(precision), device, stream, ctypes.byref(r)), "spm_sparse_codegen_real")
    return _to_perm(r, False)


def codegen_source_real(a, precision="kahan", register_efficient: bool = True) -> str:
    a = _square(a, np.float64)
    opt = _lib.SPM_CG_REGISTER_EFFICIENT if register_efficient else 0
    need = ctypes.c_size_t()
    p = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    check(lib().spm_codegen_source_real(p, a.shape[0], opt, _prec(precision), None, 0, ctypes.byref(need)),
          "spm_codegen_source_real")
    buf = ctypes.create_string_buffer(need.value)
    check(lib().spm_codegen_source_real(p, a.shape[0], opt, _prec(precision), buf, need.value, None),
          "spm_codegen_source_real")
    return buf.value.decode()


def ut_order(a) -> tuple[np.ndarray, np.ndarray]:
    """IO::UTOrder on a's pattern: b[row_perm[i], col_perm[j]] = a[i, j]."""
    a = _square(np.asarray(a, dtype=np.float64), np.float64)
    n = a.shape[0]
    rp = np.zeros(n, dtype=np.int32)
    cp = np.zeros(n, dtype=np.int32)
    check(lib().spm_ut_order(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                             rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             cp.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "spm_ut_order")
    return rp, cp


def apply_permutations(a, row_perm, col_perm) -> np.ndarray:
    a = np.asarray(a)
    b = np.empty_like(a)
    b[np.ix_(row_perm, col_perm)] = a
    return b


class CodegenPlan(DensePlan):
    """A generated SpaRyser kernel for one integer matrix (compiled once)."""

    def __init__(self, a, register_efficient: bool = True, device: int = 0):
        a = _int_square(a)
        self.complex = False
        self._h = ctypes.c_void_p()
        opt = _lib.SPM_CG_REGISTER_EFFICIENT if register_efficient else 0
        check(lib().spm_plan_codegen(a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), a.shape[0], opt, device,
                                     ctypes.byref(self._h)), "spm_plan_codegen")
        self.n = a.shape[0]

    def codegen_info(self) -> dict:
        sec, cached, h = ctypes.c_double(), ctypes.c_int(), ctypes.c_uint64()
        check(lib().spm_plan_codegen_info(self._h, ctypes.byref(sec), ctypes.byref(cached), ctypes.byref(h)))
        return {"compile_seconds": sec.value, "cached": bool(cached.value), "hash": h.value}


class CodegenRealPlan(DensePlan):
    """A generated SpaRyser kernel for one real matrix (compiled once)."""

    def __init__(self, a, precision="kahan", register_efficient: bool = True, device: int = 0):
        a = _square(a, np.float64)
        self.complex = False
        self._h = ctypes.c_void_p()
        opt = _lib.SPM_CG_REGISTER_EFFICIENT if register_efficient else 0
        check(lib().spm_plan_codegen_real(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0], opt,
                                          _prec(precision), device, ctypes.byref(self._h)), "spm_plan_codegen_real")
        self.n = a.shape[0]

    codegen_info = CodegenPlan.codegen_info


@dataclass
class Decomposition:
    components: int
    structurally_zero: bool
    avg_n: float
    avg_nnz: float
    leaves: list  # [(component, ndarray)]


def decompose(a, options: int = _lib.SPM_DECOMP_DEFAULT) -> Decomposition:
    """The reference's preprocessing (DM pruning, components, compressions)
    on the host: perm(a) = prod over components of sum of leaf permanents."""
    a = _square(a, np.float64)
    h = ctypes.c_void_p()
    check(lib().spm_decompose(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0], options,
                              ctypes.byref(h)), "spm_decompose")
    try:
        comps, nl, z = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        an, az = ctypes.c_double(), ctypes.c_double()
        check(lib().spm_decomp_stats(h, ctypes.byref(comps), ctypes.byref(nl), ctypes.byref(an), ctypes.byref(az),
                                     ctypes.byref(z)))
        leaves = []
        for k in range(nl.value):
            c, n = ctypes.c_int(), ctypes.c_int()
            check(lib().spm_decomp_leaf(h, k, ctypes.byref(c), ctypes.byref(n), None))
            m = np.zeros((n.value, n.value))
            check(lib().spm_decomp_leaf(h, k, None, None, m.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
            leaves.append((c.value, m))
        return Decomposition(comps.value, bool(z.value), an.value, az.value, leaves)
    finally:
        lib().spm_decomp_destroy(h)


def permanent_decomposed(a, precision="kahan", options: int = _lib.SPM_DECOMP_DEFAULT, device: int = 0,
                         stream: int | None = None, scaling: tuple[int, float] | None = None) -> Permanent:
    """perm(a) through the decomposition, each leaf on the GPU (exact for
    integer matrices while 128 bits hold the result).  scaling =
    (iterations, scale_into) balances every leaf first, as the reference's
    scaling=true does (spm_permanent_decomposed_scaled)."""
    a = _square(a, np.float64)
    r = SpmResult()
    ap = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    if scaling is None:
        check(lib().spm_permanent_decomposed(ap, a.shape[0], options, _prec(precision), device, stream,
                                             ctypes.byref(r)), "spm_permanent_decomposed")
    else:
        it, into = scaling
        check(lib().spm_permanent_decomposed_scaled(ap, a.shape[0], options, _prec(precision), int(it), float(into),
                                                    device, stream, ctypes.byref(r)),
              "spm_permanent_decomposed_scaled")
    return _to_perm(r, False)


def sinkhorn_scale(a, iterations: int = 100, target: float = 2.0) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The reference's IO::scale on one matrix (host only): (scaled matrix,
    row scales, column scales)."""
    b = _square(a, np.float64).copy()
    n = b.shape[0]
    rs, cs = np.zeros(n), np.zeros(n)
    D = ctypes.POINTER(ctypes.c_double)
    check(lib().spm_sinkhorn_scale(b.ctypes.data_as(D), n, int(iterations), float(target), rs.ctypes.data_as(D),
                                   cs.ctypes.data_as(D)), "spm_sinkhorn_scale")
    return b, rs, cs
