"""Generated SpaRyser kernels for real matrices and the automatic path
choice, on the GPU (through the C ABI).

Bars: within 1e-10 relative of the double-double oracle and within the
result's own error estimate (compensated mode); the generated kernel and the
dense walk (independent arithmetic) within the sum of their estimates;
row-scaled integer matrices within 2 ulp of prod(s_i) * (exact integer
permanent)."""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from conftest import GOLDEN_DIR, rel_err
from oracle import oracle as O
from superman_amd import api
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10


def real_sparse(n, seed):
    rng = np.random.default_rng(seed)
    return W.sparse01(n, seed) * rng.standard_normal((n, n))


def banded(n, seed, width=2):
    rng = np.random.default_rng(seed)
    a = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - width), min(n, i + width + 1)):
            a[i, j] = rng.standard_normal() * 10.0 ** rng.integers(-3, 4)
    return a


@pytest.mark.parametrize("prec", ["kahan", "dd"])
@pytest.mark.parametrize("register_efficient", [True, False])
@pytest.mark.parametrize("n", [5, 6, 9, 13, 17, 22])
def test_vs_dd_oracle(n, register_efficient, prec):
    a = real_sparse(n, 500 + n)
    hi, lo = O.dense_dd(a)
    r = api.sparse_codegen_real(a, prec, register_efficient)
    assert rel_err(r.value, hi + lo) <= TOL, (n, r.value, hi + lo)
    if prec == "kahan":
        assert abs(r.value - (hi + lo)) <= r.err_est


@pytest.mark.parametrize("n", [12, 20, 26])
def test_banded_wide_range_values(n):
    # a stiffness-like band (entries over 7 decades): many suffix levels
    a = banded(n, 60 + n)
    hi, lo = O.dense_dd(a)
    r = api.sparse_codegen_real(a)
    assert abs(r.value - (hi + lo)) <= r.err_est
    assert rel_err(r.value, hi + lo) <= TOL


def test_dense_real_matrix_through_generator():
    # every row touched by column 0: no suffix levels, the whole product per step
    a = W.dense_normal(14, 14)
    hi, lo = O.dense_dd(a)
    assert rel_err(api.sparse_codegen_real(a).value, hi + lo) <= TOL


def test_order_36_vs_dense_walk():
    a = real_sparse(36, 36)
    with api.CodegenRealPlan(a) as p:
        g = p.run()
        info = p.codegen_info()
    d = api.dense_f64(a)
    assert abs(g.value - d.value) <= g.err_est + d.err_est
    assert rel_err(g.value, d.value) <= TOL
    assert info["compile_seconds"] > 0


def test_plan_cache_and_determinism():
    a = real_sparse(24, 7)
    with api.CodegenRealPlan(a) as p1:
        r1 = p1.run()
    with api.CodegenRealPlan(a) as p2:
        r2 = p2.run()
        assert p2.codegen_info()["cached"]
    assert (r1.hi, r1.lo) == (r2.hi, r2.lo)


def _scaled_exact(scales, q):
    t, st = O.int_exact(q.astype(np.int64))
    assert st == 0
    v = Fraction(t)
    for s in scales:
        v *= Fraction(float(s))
    return float(v)


@pytest.mark.parametrize("n", [10, 18, 24])
def test_auto_row_scaled_integers(n):
    # rows s_i * (0/1 pattern): exact integer permanent times prod s_i
    q = W.sparse01(n, 80 + n)
    s = np.random.default_rng(n).uniform(0.1, 3.0, n)
    a = q * s[:, None]
    r, path = api.permanent_auto(a)
    assert path == "scaled_int"
    ref = _scaled_exact(s, q)
    assert abs(r.value - ref) <= 2 * math.ulp(ref), (r.value, ref)


def test_auto_paths_on_small_inputs():
    assert api.permanent_auto(W.sparse01(16, 1))[1] == "int"
    assert api.permanent_auto(W.dense_normal(16, 1))[1] == "dense"
    r, path = api.permanent_auto(0.91 * np.ones((12, 12)))
    assert path == "scaled_int"
    ref = float(Fraction(0.91) ** 12 * math.factorial(12))
    assert abs(r.value - ref) <= 2 * math.ulp(ref)


def test_example_sparse_mtx():
    # ExampleMatrices/sparse.mtx: 45 x 45, 483 entries of 0.91 -- one order-45
    # leaf; perm = 0.91^45 * perm(pattern) with an exact integer factor
    a = np.load(os.path.join(GOLDEN_DIR, "example_sparse.npy"), allow_pickle=False)
    with open(os.path.join(GOLDEN_DIR, "closed_form_examples.json")) as f:
        c = json.load(f)["sparse"]
    r, path = api.permanent_auto(a)
    assert path == "scaled_int"
    p, ppath = api.permanent_auto((a != 0).astype(np.float64))  # the same generated kernel, now cached
    assert ppath == "gen_int" and p.exact is not None
    ref = float(Fraction(0.91) ** 45 * p.exact)
    assert abs(r.value - ref) <= 2 * math.ulp(ref)
    assert abs(ref / p.exact - c["entry_power"]) <= 1e-15 * c["entry_power"]
