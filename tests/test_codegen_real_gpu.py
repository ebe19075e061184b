"""Generated SpaRyser kernels for real matrices and the automatic path
choice, on the GPU (through the C ABI).

Bars: within 1e-10 relative of the double-double oracle and within the
result's own error estimate (compensated mode) -- for the ill-conditioned
banded inputs (sum |term| / |perm| up to 3e10) the estimate alone; the
generated kernel and the dense walk (independent arithmetic) within the sum
of their estimates;
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


def dominant_band(n, seed, width=2):
    # a diagonally dominant band with rows scaled over 7 decades: row
    # scaling multiplies every Ryser term by the same factor, so the sum stays
    # well conditioned (sum |term| / |perm| about 1)
    rng = np.random.default_rng(seed)
    a = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - width), min(n, i + width + 1)):
            a[i, j] = rng.standard_normal() if i != j else 2.0 * width + abs(rng.standard_normal())
    return a * (10.0 ** rng.integers(-3, 4, n))[:, None]


@pytest.mark.parametrize("n", [12, 20, 26])
def test_banded_wide_range_values(n):
    # a stiffness-like band (entries over 7 decades): many suffix levels.
    # These permanents are ill conditioned (sum |term| / |perm| = 8e4, 3e10
    # and 4e9 at n = 12, 20, 26), so no FP64 Ryser sum reaches 1e-10 relative:
    # the bar is the result's own estimate, and agreement with the dense walk
    # (other roundings) within the two estimates.
    a = banded(n, 60 + n)
    hi, lo = O.dense_dd(a)
    r = api.sparse_codegen_real(a)
    assert abs(r.value - (hi + lo)) <= r.err_est
    d = api.dense_f64(a)
    assert abs(r.value - d.value) <= r.err_est + d.err_est


@pytest.mark.parametrize("n", [12, 20, 26])
def test_dominant_band_wide_range_values(n):
    a = dominant_band(n, 60 + n)
    hi, lo = O.dense_dd(a)
    r = api.sparse_codegen_real(a)
    assert r.abs_sum / abs(r.value) < 10.0
    assert rel_err(r.value, hi + lo) <= TOL
    assert abs(r.value - (hi + lo)) <= r.err_est


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
    # ExampleMatrices/sparse.mtx: 45 x 45, 483 entries of 0.91 and no row or
    # column with fewer than 5 -- one order-45 leaf.  20 of its rows have
    # even sums, so the exact walks carry them halved and 2^24 perm(pattern)
    # fits 128 bits: the automatic choice takes the exact path for the
    # row-scaled integers, perm = 0.91^45 perm(pattern) rounded once.
    # Checked: perm(pattern) exactly (about 45 s), against the generated
    # FP64 kernel of the pattern (independent arithmetic, about 20 s) within
    # its estimate.
    a = np.load(os.path.join(GOLDEN_DIR, "example_sparse.npy"), allow_pickle=False)
    with open(os.path.join(GOLDEN_DIR, "closed_form_examples.json")) as f:
        c = json.load(f)["sparse"]
    pattern = (a != 0).astype(np.int64)
    assert int(pattern.sum(0).min()) >= 5 and int(pattern.sum(1).min()) >= 5
    assert int((pattern.sum(1) % 2 == 0).sum()) == 20
    assert api.choose_path(a) == "scaled_int"
    exact = api.int_exact(pattern).exact
    g = api.sparse_codegen_real(pattern.astype(np.float64))
    assert abs(g.value - exact) <= g.err_est
    value = float(Fraction(0.91) ** 45 * exact)
    assert abs(value / exact - c["entry_power"]) <= 1e-15 * c["entry_power"]
    print(f"sparse.mtx: perm(pattern) = {exact} (FP64 generated kernel {g.value!r}, "
          f"cond {g.abs_sum / abs(g.value):.3g}); perm = {value!r}")
