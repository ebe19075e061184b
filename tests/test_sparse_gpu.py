"""Exact sparse walk over compressed columns (SPM_INT_WALK_SPARSE): bit-exact
against the int128 oracle and against the two dense integer walks, on 0/1
and small-integer sparse matrices, zero rows and columns, the config-3
generator at full size, and closed forms on both sides of the register-bank
split at order 32."""
import numpy as np
import pytest

from conftest import GOLDEN_DIR
from oracle import oracle as O
from superman_amd import _lib, api
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu


def sparse_ints(n, seed, per_row=4, vmax=3, signed=True):
    rng = np.random.default_rng(seed)
    a = np.zeros((n, n), dtype=np.int64)
    p = rng.permutation(n)
    for i in range(n):
        a[i, p[i]] = rng.integers(1, vmax + 1)
        k = min(int(rng.integers(0, per_row)), n - 1)
        if k == 0:
            continue
        cols = rng.choice([j for j in range(n) if j != p[i]], k, replace=False)
        lo = -vmax if signed else 1
        a[i, cols] = rng.integers(lo, vmax + 1, k)
    return a


def run(a, walk):
    with api.DensePlan(a, "exact", int_walk=walk) as p:
        assert p.info()["kernel"].startswith({"sparse": "int_sparse", "int32": "int_dense", "fp64": "int_small"}[walk])
        return p.run().exact


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 17, 20, 24])
def test_vs_oracle(n):
    for seed, vmax, signed in ((n, 1, False), (100 + n, 3, True)):
        a = sparse_ints(n, seed, per_row=4, vmax=vmax, signed=signed)
        ref, st = O.int_exact(a)
        assert st == 0
        assert run(a, "sparse") == ref, (n, seed)


def test_dense_input_and_zero_lines():
    a = (np.random.default_rng(3).random((12, 12)) < 0.5).astype(np.int64)
    ref, _ = O.int_exact(a)
    assert run(a, "sparse") == ref
    b = sparse_ints(14, 9)
    b[:, 6] = 0  # a zero column: the permanent is 0, the walk still runs every step
    assert run(b, "sparse") == 0
    c = sparse_ints(14, 10)
    c[4, :] = 0
    assert run(c, "sparse") == 0


@pytest.mark.parametrize("seed", [40, 41, 42])
def test_walks_agree_at_order_30(seed):
    # the decomposition leaves of the config-3 matrices are order 30
    d = api.decompose(np.load(f"{GOLDEN_DIR}/c3_sparse01_n40_s{seed}.npy", allow_pickle=False))
    for _, m in d.leaves:
        if m.shape[0] < 20:
            continue
        b = m.astype(np.int64)
        s = run(b, "sparse")
        assert s == run(b, "int32")
        try:
            assert s == run(b, "fp64")
        except _lib.SpmError as e:  # entries too large for the FP64 walk
            assert e.status == _lib.SPM_ERANGE


def test_config3_whole_matrix():
    a = W.sparse01(40, 40).astype(np.int64)
    assert run(a, "sparse") == 213647004
    r = api.int_exact(a)  # the automatic choice takes the sparse walk here
    assert r.exact == 213647004


@pytest.mark.parametrize("vmax,n", [(200, 12), (3000, 10), (1 << 26, 4)])
def test_large_entries_every_product_form(vmax, n):
    # row sums of 2^9, 2^13 and 2^26: products as int64 row pairs, exact up
    # to the 2^127 bound
    a = sparse_ints(n, 7 + n, per_row=3, vmax=vmax)
    ref, st = O.int_exact(a)
    assert st == 0
    assert run(a, "sparse") == ref
    assert run(a, "int32") == ref


def fib(k):
    x, y = 0, 1
    for _ in range(k):
        x, y = y, x + y
    return x


@pytest.mark.parametrize("n", [31, 32, 33, 34, 36, 41])
def test_tridiagonal_fibonacci(n):
    # perm of the 0/1 tridiagonal matrix: p(n) = p(n-1) + p(n-2), p(1) = 1,
    # p(2) = 2, so F(n+1).  Past order 32 the rows sit in two register
    # tuples and the columns at the split have entries in both.
    a = np.eye(n, dtype=np.int64) + np.eye(n, k=1, dtype=np.int64) + np.eye(n, k=-1, dtype=np.int64)
    assert run(a, "sparse") == fib(n + 1)


@pytest.mark.parametrize("n", [33, 35, 37, 38, 42, 43, 44, 45])
def test_identity_plus_cycle(n):
    # only the identity and the n-cycle avoid the zeros of I + P.  The
    # orders cover every size of the second register tuple that a walk of
    # a few seconds reaches (rows past 32: 1-12, then 16 at orders 45-48;
    # order 45 takes about 4 s); orders 49-63 (a 32-register second tuple)
    # build but take over a minute per walk and are not run here
    a = np.eye(n, dtype=np.int64) + np.roll(np.eye(n, dtype=np.int64), 1, axis=1)
    assert run(a, "sparse") == 2


@pytest.mark.parametrize("n", [33, 34, 36])
def test_both_tuples_long_columns(n):
    # signed entries, two 12-entry columns (continuation records with rows in
    # both register tuples); the int32 dense walk and a row-permuted copy must agree
    a = sparse_ints(n, 500 + n, per_row=4, vmax=3, signed=True)
    rng = np.random.default_rng(n)
    for j in (0, n - 1):
        rows = rng.choice(n, 12, replace=False)
        a[rows, j] = rng.integers(1, 4, 12)
    s = run(a, "sparse")
    assert s == run(a, "int32")
    p = np.random.default_rng(7).permutation(n)
    assert run(a[p], "sparse") == s


@pytest.mark.parametrize("n", [30, 39])
def test_every_term_nonzero(n):
    # I + 2P, P the n-cycle: every row sum is +-1 +- 2, never zero, so every
    # lane queues a product at every step (at order 39 a queue holds fewer
    # than 64 entries and one ballot overfills it); only the identity and
    # the full cycle survive: 1 + 2^n
    a = np.eye(n, dtype=np.int64) + 2 * np.roll(np.eye(n, dtype=np.int64), 1, axis=1)
    assert run(a, "sparse") == 1 + 2**n
