"""Pins the CPU oracle (CPU only).

The oracle is trusted only after these checks:
  * the restated reference walks reproduce the reference binary's own
    outputs (golden ref1, thread_count=1) bit for bit;
  * the double-double and exact-integer oracles agree with closed forms,
    brute force over permutations, an exact rational Ryser in Python, and
    algebraic invariants chosen so that a dropped term, a wrong sign or index
    or a transposed operand fails.
"""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

from conftest import golden_matrix, load_golden, rel_err
from oracle import oracle as O
from superman_amd import workloads as W

CASES = load_golden()


def brute_perm(a):
    """Sum over permutations (exact for ints / Fractions / complex)."""
    n = len(a)
    tot = 0
    for p in itertools.permutations(range(n)):
        t = 1
        for i in range(n):
            t *= a[i][p[i]]
        tot += t
    return tot


def ryser_fraction(a):
    """Plain Ryser inclusion-exclusion over column subsets in exact rationals
    (a different formula from the Gray-code NW variant under test)."""
    n = len(a)
    tot = Fraction(0)
    for mask in range(1, 1 << n):
        cols = [j for j in range(n) if mask >> j & 1]
        p = Fraction(1)
        for i in range(n):
            p *= sum(Fraction(a[i][j]) for j in cols)
        tot += (-1) ** len(cols) * p
    return (-1) ** n * tot


# ---------------------------------------------------------------- reference
@pytest.mark.parametrize("case", [c for c in CASES if "ref1" in c and c["kind"] == "real"], ids=lambda c: c["name"])
def test_dense_restatement_bit_exact_vs_reference_binary(case):
    a = golden_matrix(case)
    assert O.dense_ref(a, 1) == case["ref1"]


@pytest.mark.parametrize("case", [c for c in CASES if "ref1" in c and c["kind"] == "complex"], ids=lambda c: c["name"])
def test_complex_restatement_bit_exact_vs_reference_binary(case):
    a = golden_matrix(case)
    v = O.complex_ref(a, 1)
    assert [v.real, v.imag] == case["ref1"]


@pytest.mark.parametrize("case", [c for c in CASES if c["name"].startswith("sparse01_n")], ids=lambda c: c["name"])
def test_sparse_restatement_matches_reference_integer(case):
    # The reference's sparse CPU path sorts columns, then runs spNaivePerman;
    # on 0/1 inputs the value is an integer well below 2^53 here.
    a = golden_matrix(case)
    assert O.sparse_ref(a, 1) == case["ref1"] == float(int(case["exact"]))


def test_reference_dd_values_close():
    # ref1 (plain FP64 in the reference) vs the dd oracle: within the
    # reference's own rounding (observed <= 1e-11 relative on these inputs).
    for c in CASES:
        if "ref1" in c and "dd" in c:
            dd = sum(c["dd"][:2]) if c["kind"] == "real" else complex(c["dd"][0] + c["dd"][1], c["dd"][2] + c["dd"][3])
            ref = c["ref1"] if c["kind"] == "real" else complex(*c["ref1"])
            assert rel_err(ref, dd) < 1e-9, c["name"]


# ---------------------------------------------------------------- closed forms
def test_known_answers():
    assert O.int_exact(np.array([[1, 2, 3], [4, 5, 6], [7, 8, 10]]))[0] == 463
    assert O.int_exact(np.ones((12, 12), dtype=np.int64))[0] == math.factorial(12)
    hi, lo = O.dense_dd(np.full((10, 10), 0.5))
    assert hi == math.factorial(10) * 0.5 ** 10
    v = O.complex_dd(np.full((10, 10), 1j))
    assert (v[0], v[2]) == (-math.factorial(10), 0.0)


@pytest.mark.parametrize("n", range(1, 11))
def test_derangements(n):
    # perm(J - I) = number of derangements D_n
    d = [1, 0]
    for k in range(2, 12):
        d.append((k - 1) * (d[-1] + d[-2]))
    a = np.ones((n, n), dtype=np.int64) - np.eye(n, dtype=np.int64)
    assert O.int_exact(a)[0] == d[n]
    assert O.dense_dd(a.astype(float))[0] == d[n]


def test_menage_numbers():
    # perm(J - I - P) with P the cyclic shift: menage numbers U_n
    U = {3: 1, 4: 2, 5: 13, 6: 80, 7: 579, 8: 4738, 9: 43387, 10: 439792}
    for n, u in U.items():
        a = np.ones((n, n), dtype=np.int64) - np.eye(n, dtype=np.int64) - np.roll(np.eye(n, dtype=np.int64), 1, axis=1)
        assert O.int_exact(a)[0] == u, n


def test_triangular_and_permutation():
    rng = np.random.default_rng(5)
    for n in (3, 6, 11):
        t = np.triu(rng.integers(-9, 10, (n, n)))
        assert O.int_exact(t)[0] == int(np.prod(np.diag(t)))
        p = np.eye(n, dtype=np.int64)[rng.permutation(n)]
        assert O.int_exact(p)[0] == 1
        hi, lo = O.dense_dd(np.triu(rng.random((n, n))))


# ---------------------------------------------------------------- brute force
@pytest.mark.parametrize("n", range(1, 8))
def test_int_exact_vs_brute_force(n):
    rng = np.random.default_rng(100 + n)
    a = rng.integers(-5, 6, (n, n))
    assert O.int_exact(a)[0] == brute_perm(a.tolist())


@pytest.mark.parametrize("n", range(1, 8))
def test_dense_dd_vs_exact_rational(n):
    # dyadic-exact inputs: every double is an exact Fraction
    rng = np.random.default_rng(200 + n)
    a = rng.standard_normal((n, n))
    exact = brute_perm([[Fraction(v) for v in row] for row in a.tolist()])
    hi, lo = O.dense_dd(a)
    assert abs(Fraction(hi) + Fraction(lo) - exact) <= abs(exact) * Fraction(1, 10 ** 25) + Fraction(1, 10 ** 300)


@pytest.mark.parametrize("n", [9, 11, 12])
def test_dense_dd_vs_rational_ryser(n):
    rng = np.random.default_rng(300 + n)
    a = rng.standard_normal((n, n))
    exact = ryser_fraction(a.tolist())
    hi, lo = O.dense_dd(a)
    assert abs(Fraction(hi) + Fraction(lo) - exact) <= abs(exact) * Fraction(1, 10 ** 24)


@pytest.mark.parametrize("n", range(1, 7))
def test_complex_dd_vs_brute_force(n):
    rng = np.random.default_rng(400 + n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    fr = [[(Fraction(v.real), Fraction(v.imag)) for v in row] for row in a.tolist()]

    def cm(x, y):
        return (x[0] * y[0] - x[1] * y[1], x[0] * y[1] + x[1] * y[0])

    tot = (Fraction(0), Fraction(0))
    for p in itertools.permutations(range(n)):
        t = (Fraction(1), Fraction(0))
        for i in range(n):
            t = cm(t, fr[i][p[i]])
        tot = (tot[0] + t[0], tot[1] + t[1])
    d = O.complex_dd(a)
    err = abs(complex(float(Fraction(d[0]) + Fraction(d[1]) - tot[0]), float(Fraction(d[2]) + Fraction(d[3]) - tot[1])))
    assert err <= 1e-25 * abs(complex(float(tot[0]), float(tot[1])))


# ---------------------------------------------------------------- invariants
def dd_rel(x, y):
    """Relative difference of two (hi, lo) double-double values."""
    return abs((x[0] - y[0]) + (x[1] - y[1])) / abs(y[0])


def test_invariants_dd():
    rng = np.random.default_rng(7)
    a = rng.standard_normal((10, 10))
    base = O.dense_dd(a)
    assert dd_rel(O.dense_dd(a.T), base) < 1e-24
    assert dd_rel(O.dense_dd(a[rng.permutation(10)][:, rng.permutation(10)]), base) < 1e-24
    s = a.copy()
    s[3] *= 2.0  # exact scaling: every operation of the walk scales exactly
    assert O.dense_dd(s) == (2 * base[0], 2 * base[1])
    # block diagonal -> product
    b = np.zeros((12, 12))
    b[:5, :5] = a[:5, :5]
    b[5:, 5:] = rng.standard_normal((7, 7))
    p1, p2 = O.dense_dd(b[:5, :5]), O.dense_dd(b[5:, 5:])
    prod = Fraction(p1[0]) + Fraction(p1[1])
    prod *= Fraction(p2[0]) + Fraction(p2[1])
    pb = O.dense_dd(b)
    assert abs(Fraction(pb[0]) + Fraction(pb[1]) - prod) <= abs(prod) * Fraction(1, 10 ** 24)
    # conj(perm A) = perm(conj A)
    c = W.complex_gauss(8, 3)
    d1, d2 = O.complex_dd(c), O.complex_dd(np.conj(c))
    assert dd_rel((d2[0], d2[1]), (d1[0], d1[1])) < 1e-24
    assert dd_rel((d2[2], d2[3]), (-d1[2], -d1[3])) < 1e-24


def test_int_divisibility_self_check_and_bound():
    a = W.sparse01(20, 3).astype(np.int64)
    v, st = O.int_exact(a)
    assert st == 0 and v > 0
    assert O.int_log2_bound(a) < 126


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_oracles_independent_of_thread_count(threads):
    a = W.dense_normal(14, 14)
    assert dd_rel(O.dense_dd(a, threads), O.dense_dd(a, 1)) < 1e-26
    b = (np.random.default_rng(1).random((14, 14)) < 0.5).astype(np.int64)
    assert O.int_exact(b, threads) == O.int_exact(b, 1)


def test_golden_exact_values_reproduce():
    for c in CASES:
        if "exact" in c and c["n"] <= 24:
            v, st = O.int_exact(golden_matrix(c).astype(np.int64))
            assert str(v) == c["exact"] and st == c.get("exact_status", 0), c["name"]


@pytest.mark.parametrize("n", [5, 12, 18])
def test_dense_ref_sample_full_range_is_the_whole_walk(n):
    # the CPU-baseline sample over every Gray index [1, 2^(n-1)) is the
    # reference walk's sum without the k=0 term and the final factor: with
    # one thread both run the same additions in the same order
    a = W.dense_normal(n, 300 + n)
    x = [a[i, n - 1] - sum(a[i, j] for j in range(n)) / 2 for i in range(n)]  # left-to-right row sums
    p0 = 1.0
    for v in x:
        p0 *= v
    part = O.dense_ref_sample(a, 1, (1 << (n - 1)) - 1)
    assert (4 * (n % 2) - 2) * (p0 + part) == O.dense_ref(a, 1)
    # thread splits of the sample change the addition order only
    assert rel_err(O.dense_ref_sample(a, 4, (1 << (n - 1)) - 1), part) <= 1e-12


def test_dense_ref_sample_prefix_vs_direct_subset_sum():
    # a prefix of the walk against the terms formed from scratch: k = 1..count,
    # S = Gray(k) = k ^ (k >> 1), term (-1)^k prod_i (x_i + sum_{j in S} a_ij)
    n, count = 11, 300
    a = W.dense_uniform(n, 77)
    x = a[:, n - 1] - a.sum(axis=1) / 2
    direct = 0.0
    for k in range(1, count + 1):
        g = k ^ (k >> 1)
        cols = [j for j in range(n - 1) if (g >> j) & 1]
        direct += (-1) ** k * np.prod(x + a[:, cols].sum(axis=1))
    assert rel_err(O.dense_ref_sample(a, 1, count), direct) <= 1e-12
    assert rel_err(O.dense_ref_sample(a, 3, count), direct) <= 1e-12


def test_complex_ref_sample_full_range_is_the_whole_walk():
    n = 12
    z = W.complex_gauss(n, 91)
    x = z[:, n - 1] - z.sum(axis=1) / 2
    p0 = complex(1.0, 0.0)
    for v in x:
        p0 = complex(p0.real * v.real - p0.imag * v.imag, p0.real * v.imag + p0.imag * v.real)
    part = O.complex_ref_sample(z, 1, (1 << (n - 1)) - 1)
    whole = O.complex_ref(z, 1)
    f = 4 * (n % 2) - 2
    assert rel_err(f * (p0 + part), whole) <= 1e-13
    # a prefix against terms formed from scratch
    count = 200
    direct = 0j
    for k in range(1, count + 1):
        g = k ^ (k >> 1)
        cols = [j for j in range(n - 1) if (g >> j) & 1]
        direct += (-1) ** k * np.prod(x + z[:, cols].sum(axis=1))
    assert rel_err(O.complex_ref_sample(z, 2, count), direct) <= 1e-12
