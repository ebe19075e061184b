"""FP32-product / FP64-accumulate variant for nonnegative matrices
(SPM_PREC_F32_PRODUCT): every result within its stated error estimate of the
double-double oracle, exact power-of-two row-scaling behaviour, and the
measured error next to the bound."""
import math

import numpy as np
import pytest

from conftest import rel_err
from oracle import oracle as O
import superman_amd as S
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu


def f32(a):
    return S.dense_f64(a, precision="f32_product")


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 13, 16, 20, 24])
def test_within_bound_of_oracle(n):
    a = W.dense_uniform(n, 100 + n)
    hi, lo = O.dense_dd(a)
    r = f32(a)
    err = abs(r.value - (hi + lo))
    print(f"n={n} rel err {err / abs(hi):.2e} estimate {r.err_est / abs(hi):.2e}")
    assert err <= r.err_est
    assert r.err_est / abs(hi) < 1e-2


def test_closed_form_constant_matrix():
    n, c = 12, 0.37
    r = f32(np.full((n, n), c))
    exact = c ** n * math.factorial(n)
    assert abs(r.value - exact) <= r.err_est
    assert rel_err(r.value, exact) < 1e-4


def test_row_scaling_by_powers_of_two_is_exact():
    a = W.dense_uniform(14, 7)
    base = f32(a)
    b = a.copy()
    b[3] *= 2.0 ** 40
    b[9] *= 2.0 ** -17
    scaled = f32(b)
    assert scaled.value == math.ldexp(base.value, 23)
    big = f32(a * 2.0 ** 33)  # entries ~1e10: far outside FP32's product range unscaled
    assert big.value == math.ldexp(base.value, 33 * 14)


def test_zero_row_and_identity():
    a = W.dense_uniform(9, 3)
    a[4] = 0.0
    assert f32(a).value == 0.0
    assert f32(np.eye(10)).value == 1.0


def test_sparse_nonnegative_integer_matrix():
    a = W.sparse01(20, 5)
    exact, st = O.int_exact(a.astype(np.int64))
    assert st == 0
    r = f32(a)
    assert abs(r.value - exact) <= r.err_est
