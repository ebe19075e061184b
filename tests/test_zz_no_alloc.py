"""spm_dense_f64_noalloc (the dense FP64 permanent with no allocation per
call) against spm_dense_f64 through the C ABI on the GPU.

Bar: bit-identical records (hi, lo, abs_sum, err_est, value) -- both run the
same kernels over the same geometry on the same host preparation.  Each
case makes two calls in a row on the shared per-device array, the second
with a smaller matrix, so a value left over from the first call would show.
"""
import ctypes

import numpy as np
import pytest

from superman_amd import _lib
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu


def call(name, a, precision):
    a = np.ascontiguousarray(a, dtype=np.float64)
    r = _lib.SpmResult()
    st = getattr(_lib.lib(), name)(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0], precision, 0,
                                   None, ctypes.byref(r))
    return st, r


def same(r, q):
    return (r.re, r.hi_re, r.lo_re, r.abs_sum, r.err_est, r.n, r.subsets, r.launches) == \
           (q.re, q.hi_re, q.lo_re, q.abs_sum, q.err_est, q.n, q.subsets, q.launches)


@pytest.mark.parametrize("precision", [_lib.SPM_PREC_COMPENSATED, _lib.SPM_PREC_PLAIN, _lib.SPM_PREC_F32_PRODUCT])
@pytest.mark.parametrize("orders", [(30, 20), (24, 5), (13, 1)])
def test_two_calls_match_dense_f64(orders, precision):
    gen = W.dense_uniform if precision == _lib.SPM_PREC_F32_PRODUCT else W.dense_normal
    for n in orders:
        a = gen(n, 700 + n)
        st0, r0 = call("spm_dense_f64", a, precision)
        st1, r1 = call("spm_dense_f64_noalloc", a, precision)
        assert st0 == _lib.SPM_OK and st1 == _lib.SPM_OK, (n, st0, st1)
        assert same(r0, r1), (n, r0.re, r1.re, r0.lo_re, r1.lo_re)
        assert r1.kernel_ms >= 0.0


def test_order_42_two_launches():
    # a run long enough to be split into launches that accumulate into the
    # per-block records (about 5 s each call)
    a = W.dense_normal(42, 42)
    st0, r0 = call("spm_dense_f64", a, _lib.SPM_PREC_PLAIN)
    st1, r1 = call("spm_dense_f64_noalloc", a, _lib.SPM_PREC_PLAIN)
    assert st0 == st1 == _lib.SPM_OK
    assert r1.launches >= 2
    assert same(r0, r1)


def test_rejects_out_of_range_arguments():
    a = np.ones((64, 64))
    st, r = call("spm_dense_f64_noalloc", a, _lib.SPM_PREC_COMPENSATED)
    assert st == _lib.SPM_EINVAL and r.status == _lib.SPM_EINVAL
    st, r = call("spm_dense_f64_noalloc", np.ones((4, 4)), 9)
    assert st == _lib.SPM_EINVAL
    st, r = call("spm_dense_f64_noalloc", -np.ones((4, 4)), _lib.SPM_PREC_F32_PRODUCT)
    assert st == _lib.SPM_EINVAL
