"""The automatic path choice (spm_choose_path, host only): MI355X's cost
model pinned on a table of shapes, checked without a GPU.  The rows follow
times measured on MI355X at orders 16-40 (every path, first call with the
hiprtc compile and cached) and the kernels' per-subset rates at order 40."""
import math

import numpy as np
import pytest

from superman_amd import api


def pattern(n, per_row, seed):
    """0/1, a seeded permutation plus per_row - 1 further entries per row."""
    if per_row >= n:
        return np.ones((n, n))
    rng = np.random.default_rng(seed)
    a = np.zeros((n, n))
    perm = rng.permutation(n)
    for i in range(n):
        a[i, perm[i]] = 1
        others = rng.choice([j for j in range(n) if j != perm[i]], size=per_row - 1, replace=False)
        a[i, others] = 1
    return a


def real_pattern(n, per_row, seed):
    return pattern(n, per_row, seed) * np.random.default_rng(seed + 1).standard_normal((n, n))


TABLE = [
    # below order ~30 a call is host-overhead bound: no compile pays
    (16, 4, "int01", "int"),
    (16, 4, "real", "dense"),
    (36, 4, "real", "dense"),           # 6 ms generated + 0.24 s compile against 71 ms dense
    # order 40, 4 per row: the compressed-column walk (0.23 s) beats a
    # 0.32 s compile + 0.05 s generated kernel; real entries: the generated
    # FP64 kernel (0.09 s + 0.25 s compile) beats the dense walk (1.26 s)
    (40, 4, "int01", "int"),
    (40, 4, "real", "gen_real"),
    (40, 40, "real", "dense"),          # a dense pattern: suffix products do not help
    # order 44: the generated integer kernel pays for its compile at 4
    # entries per row ...
    (44, 4, "int01", "gen_int"),
    # ... but not at 8: the compressed-column walk is faster than the
    # generated kernel there (order 40: 565 ms against 661 ms + a 2.0 s
    # compile)
    (44, 8, "int01", "int"),
    # ... unless every row sum is odd: no row is ever zero, every term is
    # multiplied out, and the exact FP64 walk wins
    (44, 3, "int01", "int"),
    (44, 44, "real", "dense"),
]


@pytest.mark.parametrize("n,per_row,kind,expected", TABLE)
def test_choice_table(n, per_row, kind, expected):
    gen = pattern if kind == "int01" else real_pattern
    a = gen(n, per_row, 1000 * n + per_row)
    assert api.choose_path(a) == expected


def test_row_scaled_integers_take_the_exact_path():
    assert api.choose_path(0.91 * pattern(20, 4, 5)) == "scaled_int"
    s = np.random.default_rng(2).uniform(0.5, 2.0, 24)
    assert api.choose_path(pattern(24, 4, 6) * s[:, None]) == "scaled_int"


def test_odd_row_sums_never_skip():
    # the estimate sees the parity: equal patterns up to one entry per row
    # cost very differently for the generated integer kernel
    odd = pattern(44, 3, 44003)
    even = pattern(44, 4, 44004)
    assert api.path_estimate(odd, "gen_int") > 5 * api.path_estimate(even, "gen_int")


def test_estimates_grow_with_the_order():
    for path in ("dense", "int"):
        t = [api.path_estimate(pattern(n, 4, n), path) for n in (30, 34, 38)]
        assert t[0] < t[1] < t[2]
    assert math.isclose(api.path_estimate(np.ones((40, 40)), "dense"), 3e-5 + 2 ** 39 * 80 / 3.5e13, rel_tol=1e-12)
