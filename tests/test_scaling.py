"""The reference's scaling=true (IO::scale, IO/IO.h:109-161, applied per
leaf by DecomposePerman, CommonFiles/DecomposePerman.h:97-107, 189-193),
checked without a GPU: the library's balancing against the oracle's
restatement (bit-identical scales and scaled matrix), the restatement
against what the mathematics fixes, and argument checks of the C ABI."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from superman_amd import _lib, api
from superman_amd import workloads as W


@pytest.mark.parametrize("gen,n,seed,it,target", [
    ("dense_uniform", 8, 1, 100, 2.0),
    ("dense_uniform", 13, 2, 5, 3.0),
    ("sparse01", 12, 3, 100, 2.0),
    ("dense_normal", 9, 4, 20, 2.0),
])
def test_library_equals_restatement(gen, n, seed, it, target):
    a = getattr(W, gen)(n, seed)
    b, r, c = api.sinkhorn_scale(a, it, target)
    ob, orr, oc = O.sinkhorn_scale(a, it, target)
    assert np.array_equal(r, orr) and np.array_equal(c, oc)
    assert np.array_equal(b, ob)


def test_restatement_closed_form_all_ones():
    # J_4, one round towards 2: row sums 4 -> r_i = 1/2; then column sums
    # 4 * 1/2 = 2 -> c_j = 1; every scaled entry 1/2
    b, r, c = O.sinkhorn_scale(np.ones((4, 4)), 1, 2.0)
    assert np.all(r == 0.5) and np.all(c == 1.0) and np.all(b == 0.5)


def test_restatement_replaces_scales_like_the_reference():
    # IO::scale sets each scale to target / (line sum at the current
    # scales), replacing it rather than multiplying it as textbook Sinkhorn
    # would: on J_4 towards 2 the second round gives r = 1, c = 1/2 (textbook:
    # r = 1/2, c = 1 again).  The scaled entries are 1/2 either way.
    b, r, c = O.sinkhorn_scale(np.ones((4, 4)), 2, 2.0)
    assert np.all(r == 1.0) and np.all(c == 0.5) and np.all(b == 0.5)


def test_restatement_scales_out_of_the_permanent():
    # b = diag(r) a diag(c) for any scales, so perm(b) = prod(r) prod(c) perm(a)
    a = W.dense_uniform(10, 7) + 0.1
    b, r, c = O.sinkhorn_scale(a, 100, 2.0)
    assert np.allclose(b, r[:, None] * a * c[None, :], rtol=1e-15, atol=0)
    hb, lb = O.dense_dd(b)
    ha, la = O.dense_dd(a)
    assert abs((hb + lb) / (np.prod(r) * np.prod(c)) - (ha + la)) <= 1e-12 * abs(ha + la)


def test_zero_row_keeps_its_scale():
    a = W.dense_uniform(6, 5)
    a[2, :] = 0.0
    b, r, c = api.sinkhorn_scale(a, 10, 2.0)
    assert r[2] == 1.0 and np.all(b[2] == 0.0)


def test_scaled_entry_points_reject_bad_arguments_without_gpu():
    L = _lib.lib()
    a = np.ones((5, 5))
    D = ctypes.POINTER(ctypes.c_double)
    rs, cs = np.zeros(5), np.zeros(5)
    assert L.spm_sinkhorn_scale(a.ctypes.data_as(D), 5, 10, 0.0, rs.ctypes.data_as(D),
                                cs.ctypes.data_as(D)) == _lib.SPM_EINVAL
    assert L.spm_sinkhorn_scale(a.ctypes.data_as(D), 5, -1, 2.0, rs.ctypes.data_as(D),
                                cs.ctypes.data_as(D)) == _lib.SPM_EINVAL
    r = _lib.SpmResult()
    assert L.spm_permanent_decomposed_scaled(a.ctypes.data_as(D), 5, _lib.SPM_DECOMP_DEFAULT, 1, -3, 2.0, 0, None,
                                             ctypes.byref(r)) == _lib.SPM_EINVAL
