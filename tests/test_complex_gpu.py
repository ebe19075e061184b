"""Dense complex128 Gray-code Ryser on the GPU vs the double-double oracle
(through the C ABI).  Bar: relative error <= 1e-10, and the
compensated mode inside its own stated error estimate."""
import math

import numpy as np
import pytest

from conftest import golden_matrix, load_golden, rel_err
from oracle import oracle as O
from superman_amd import api
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10


def dd_value(a):
    d = O.complex_dd(a)
    return complex(d[0] + d[1], d[2] + d[3])


@pytest.mark.parametrize("prec", ["dd", "kahan"])
@pytest.mark.parametrize("n", list(range(1, 21)))
def test_gauss_vs_dd_oracle(n, prec):
    a = W.complex_gauss(n, 3000 + n)
    ref = dd_value(a)
    r = api.dense_c128(a, prec)
    assert rel_err(r.value, ref) <= TOL, (n, r.value, ref)
    if prec == "kahan":
        assert abs(r.value - ref) <= r.err_est, (n, abs(r.value - ref), r.err_est)


@pytest.mark.parametrize("case", [c for c in load_golden() if c["kind"] == "complex" and "dd" in c],
                         ids=lambda c: c["name"])
def test_golden_complex(case):
    a = golden_matrix(case)
    d = case["dd"]
    ref = complex(d[0] + d[1], d[2] + d[3])
    r = api.dense_c128(a, "kahan")
    assert rel_err(r.value, ref) <= TOL, (r.value, ref)
    assert abs(r.value - ref) <= r.err_est
    if "ref1" in case:  # the reference's own (uncompensated) output lies within its rounding
        assert rel_err(complex(*case["ref1"]), r.value) <= 1e-9


def test_closed_forms():
    # perm(c J_n) = c^n n!  (all-i: the reference's complex.mtx family)
    for n in (6, 10, 12):
        r = api.dense_c128(np.full((n, n), 1j))
        assert rel_err(r.value, (1j) ** n * math.factorial(n)) <= 1e-13
    # diagonal: product of the diagonal
    d = W.complex_gauss(18, 5).diagonal().copy()
    r = api.dense_c128(np.diag(d))
    assert rel_err(r.value, np.prod(d)) <= 1e-13


def test_conjugate_and_real_agree():
    a = W.complex_gauss(17, 17)
    r1 = api.dense_c128(a)
    r2 = api.dense_c128(np.conj(a))
    assert rel_err(r2.value, np.conj(r1.value)) <= TOL
    b = W.dense_normal(19, 19)
    rc = api.dense_c128(b.astype(np.complex128))
    rr = api.dense_f64(b)
    assert rel_err(rc.value, rr.value) <= TOL and abs(rc.value.imag) <= 1e-12 * abs(rr.value)


def test_haar_n32_plain_vs_compensated(golden):
    c = golden["c4_haar_n32_s32"]
    a = golden_matrix(c)
    comp = api.dense_c128(a, "kahan")
    plain = api.dense_c128(a, "dd")
    assert abs(plain.value - comp.value) <= comp.err_est
    if "dd" in c:
        d = c["dd"]
        ref = complex(d[0] + d[1], d[2] + d[3])
        assert rel_err(comp.value, ref) <= TOL


@pytest.mark.parametrize("n", [33, 36])
def test_orders_with_three_column_parts(n):
    # orders 33..48 take each column in three parts of at most 32 doubles
    r = api.dense_c128(np.full((n, n), 1j))
    exact = (1j) ** n * math.factorial(n)
    assert abs(r.value - exact) <= r.err_est and rel_err(r.value, exact) <= TOL
    b = W.dense_normal(n, n)
    rc = api.dense_c128(b.astype(np.complex128))
    rr = api.dense_f64(b)
    assert abs(rc.value - rr.value) <= rc.err_est + rr.err_est
    assert abs(rc.value.imag) == 0.0


@pytest.mark.parametrize("n", [22, 23, 26])
def test_small_order_chunks_invariance(n):
    # orders whose chunk length comes from the small-order rule (see the
    # real walk's test): transpose and permutations agree within the estimates
    a = W.complex_gauss(n, 400 + n)
    rng = np.random.default_rng(n)
    b = a[rng.permutation(n)][:, rng.permutation(n)]
    r, rt, rp = api.dense_c128(a), api.dense_c128(np.ascontiguousarray(a.T)), api.dense_c128(b)
    for s in (rt, rp):
        assert abs(s.value - r.value) <= r.err_est + s.err_est
