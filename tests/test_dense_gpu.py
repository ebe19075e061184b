"""Dense real FP64 Gray-code Ryser on the GPU vs the oracle (through the C ABI).

Bar: relative error <= 1e-10 against the double-double oracle;
the compensated mode must also sit inside its own stated error estimate.
At n = 40 (no oracle) the checks are size-independent properties."""
import numpy as np
import pytest

from conftest import golden_matrix, load_golden, rel_err
from oracle import oracle as O
from superman_amd import api
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10


def dd_value(a):
    hi, lo = O.dense_dd(a)
    return hi + lo


@pytest.mark.parametrize("prec", ["dd", "kahan"])
@pytest.mark.parametrize("n", list(range(1, 23)))
def test_uniform_vs_dd_oracle(n, prec):
    a = W.dense_uniform(n, 1000 + n)
    ref = dd_value(a)
    r = api.dense_f64(a, prec)
    assert rel_err(r.value, ref) <= TOL, (n, r.value, ref)


@pytest.mark.parametrize("prec", ["dd", "kahan"])
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 9, 12, 15, 18, 21, 24])
def test_normal_vs_dd_oracle(n, prec):
    a = W.dense_normal(n, 2000 + n)
    ref = dd_value(a)
    r = api.dense_f64(a, prec)
    assert rel_err(r.value, ref) <= TOL, (n, r.value, ref)
    if prec == "kahan":
        assert abs(r.value - ref) <= r.err_est, (n, abs(r.value - ref), r.err_est)


@pytest.mark.parametrize("case", [c for c in load_golden() if c["kind"] == "real" and "dd" in c],
                         ids=lambda c: c["name"])
def test_golden_real(case):
    a = golden_matrix(case)
    ref = case["dd"][0] + case["dd"][1]
    for prec in ("dd", "kahan"):
        r = api.dense_f64(a, prec)
        assert rel_err(r.value, ref) <= TOL, (prec, r.value, ref)
    # the reference's own output lies within its rounding of ours
    if "ref1" in case:
        assert rel_err(case["ref1"], r.value) <= 1e-9


def test_integer_valued_known_answers():
    assert api.dense_f64([[1, 2, 3], [4, 5, 6], [7, 8, 10]]).value == 463.0
    assert api.dense_f64(np.ones((12, 12))).value == 479001600.0
    assert api.dense_f64(np.eye(20)).value == 1.0
    d = [1, 0]
    for k in range(2, 16):
        d.append((k - 1) * (d[-1] + d[-2]))
    for n in (5, 9, 15):
        assert api.dense_f64(np.ones((n, n)) - np.eye(n)).value == d[n]


def test_plan_reuse_deterministic():
    a = W.dense_normal(26, 26)
    with api.DensePlan(a, "kahan") as p:
        r1, r2 = p.run(), p.run()
    assert (r1.hi, r1.lo) == (r2.hi, r2.lo)


@pytest.mark.parametrize("n", [28, 30])
def test_oracle_sizes_normal(n, golden):
    c = golden[f"dense_normal_n{n}_s{n}"]
    a = golden_matrix(c)
    ref = c["dd"][0] + c["dd"][1]
    r = api.dense_f64(a, "kahan")
    assert rel_err(r.value, ref) <= TOL
    assert abs(r.value - ref) <= r.err_est


def test_n40_properties(golden):
    a = golden_matrix(golden["c2_dense_normal_n40_s40"])
    comp = api.dense_f64(a, "kahan")
    plain = api.dense_f64(a, "dd")
    # plain vs compensated agree within the stated estimate
    assert abs(plain.value - comp.value) <= comp.err_est
    assert rel_err(plain.value, comp.value) <= TOL
    # exact row scaling by 2: every operation scales exactly -> bit-exact
    s = a.copy()
    s[17] *= 2.0
    comp2 = api.dense_f64(s, "kahan")
    assert (comp2.hi, comp2.lo) == (2 * comp.hi, 2 * comp.lo)
    # transpose: different arithmetic, same permanent
    t = api.dense_f64(np.ascontiguousarray(a.T), "kahan")
    assert rel_err(t.value, comp.value) <= TOL


@pytest.mark.parametrize("n", [23, 25, 26, 27])
def test_small_order_chunks_invariance(n):
    # orders whose chunk length comes from the small-order rule: transpose
    # and row/column permutations walk the columns in other orders and must
    # agree within the stated estimates
    a = W.dense_normal(n, 300 + n)
    rng = np.random.default_rng(n)
    b = a[rng.permutation(n)][:, rng.permutation(n)]
    r, rt, rp = api.dense_f64(a, "kahan"), api.dense_f64(np.ascontiguousarray(a.T), "kahan"), api.dense_f64(b, "kahan")
    for s in (rt, rp):
        assert abs(s.value - r.value) <= r.err_est + s.err_est
