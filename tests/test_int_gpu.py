"""Exact integer Gray-code Ryser on the GPU (through the C ABI).

Bar: bit-exact.  Checked against the int128 CPU oracle, the reference's own
outputs (tests/golden: the sparse 0/1 n=40 inputs, which the reference
computes through its decomposition), and closed forms."""
import math

import numpy as np
import pytest

from conftest import golden_matrix, load_golden
from oracle import oracle as O
from superman_amd import _lib, api
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", list(range(1, 17)))
def test_random_small_ints_vs_oracle(n):
    a = np.random.default_rng(500 + n).integers(-5, 6, (n, n))
    ref, st = O.int_exact(a)
    assert st == 0
    r = api.int_exact(a)
    assert r.exact == ref, (n, r.exact, ref)
    assert r.value == float(ref)


@pytest.mark.parametrize("case", [c for c in load_golden() if "exact" in c], ids=lambda c: c["name"])
def test_golden_exact(case):
    a = golden_matrix(case)
    r = api.int_exact(a)
    assert r.exact == int(case["exact"])
    if "known" in case:
        assert r.exact == int(case["known"])


@pytest.mark.parametrize("case", [c for c in load_golden() if c["name"].startswith("c3_sparse01_n40")],
                         ids=lambda c: c["name"])
def test_sparse01_n40_matches_reference(case):
    # the reference's value (computed through its DM/component/degree
    # decomposition) is an exact integer below 2^53 for these inputs
    a = golden_matrix(case)
    ref = case["ref8"]
    assert ref == int(ref)
    r = api.int_exact(a)
    assert r.exact == int(ref), (r.exact, ref)


def test_closed_forms():
    d = [1, 0]
    for k in range(2, 30):
        d.append((k - 1) * (d[-1] + d[-2]))
    for n in (2, 7, 13, 20, 24):
        assert api.int_exact(np.ones((n, n), dtype=np.int64) - np.eye(n, dtype=np.int64)).exact == d[n]
    U = {5: 13, 8: 4738, 10: 439792}
    for n, u in U.items():
        a = np.ones((n, n), dtype=np.int64) - np.eye(n, dtype=np.int64) - np.roll(np.eye(n, dtype=np.int64), 1, 1)
        assert api.int_exact(a).exact == u
    for n in (12, 16, 20):
        assert api.int_exact(np.ones((n, n), dtype=np.int64)).exact == math.factorial(n)
    # triangular: product of the diagonal (size chosen so the row-sum bound
    # 2^(n-1) prod_i sum_j |a_ij| stays below 2^126)
    t = np.triu(np.random.default_rng(9).integers(-3, 4, (16, 16)))
    np.fill_diagonal(t, np.where(np.diag(t) == 0, 2, np.diag(t)))
    assert api.int_exact(t).exact == int(np.prod(np.diag(t).astype(object)))


def test_dense01_half_density_n28(golden):
    c = golden["int01_d50_n28_s28"]
    r = api.int_exact(golden_matrix(c))
    assert r.exact == int(c["exact"])


def test_plan_reuse_and_permutation_invariance():
    a = W.sparse01(30, 7).astype(np.int64)
    with api.DensePlan(a, "exact") as p:
        r1, r2 = p.run(), p.run()
    assert r1.exact == r2.exact == 2128528
    rng = np.random.default_rng(1)
    b = a[rng.permutation(30)][:, rng.permutation(30)]
    assert api.int_exact(b).exact == 2128528
    assert api.int_exact(np.ascontiguousarray(a.T)).exact == 2128528


def test_range_refused():
    a = np.random.default_rng(2).integers(-9, 10, (40, 40))
    with pytest.raises(_lib.SpmError) as e:
        api.int_exact(a)
    assert e.value.status == _lib.SPM_ERANGE


def test_auto_dispatch_uses_exact_path():
    a = (np.random.default_rng(24).random((24, 24)) < 0.5).astype(np.float64)
    r = api.permanent(a)
    assert r.exact is not None and r.exact == O.int_exact(a.astype(np.int64))[0]


def _walk(a, walk):
    with api.DensePlan(np.asarray(a, dtype=np.int64), "exact", int_walk=walk) as p:
        return p.run().exact, p.info()["kernel"]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 13, 17, 22, 24, 26, 28])
def test_int_walks_agree(n):
    # entries in {-1, 0, 1}: absolute row sums <= n < 2^(bit width of n),
    # so both the int32 walk and the FP64 walk apply; about 5 nonzeros per
    # row keeps 2^(n-1) prod_i sum_j |a_ij| below the 2^126 exactness bound
    rng = np.random.default_rng(900 + n)
    a = rng.choice([-1, 1], (n, n)) * (rng.random((n, n)) < min(1.0, 5.0 / n))
    ref, st = O.int_exact(a)
    assert st == 0
    v32, k32 = _walk(a, "int32")
    v64, k64 = _walk(a, "fp64")
    assert k32.startswith("int_dense_kernel") and k64.startswith("int_small_kernel")
    assert v32 == v64 == ref, (n, v32, v64, ref)


@pytest.mark.parametrize("case", [c for c in load_golden() if c["name"].startswith(("int01_d50", "sparse01_n2"))],
                         ids=lambda c: c["name"])
def test_int_walks_agree_golden(case):
    a = golden_matrix(case)
    assert _walk(a, "int32")[0] == _walk(a, "fp64")[0] == int(case["exact"])


def test_fp64_walk_sparse01_n40_matches_reference(golden):
    c = golden["c3_sparse01_n40_s40"]
    v, k = _walk(golden_matrix(c), "fp64")
    assert k == "int_small_kernel<40>" and v == int(c["ref8"])


def _blocks_of_ones(n_block=10, blocks=4, seed=3):
    rng = np.random.default_rng(seed)
    n = blocks * n_block
    j = np.zeros((n, n), dtype=np.int64)
    for b in range(blocks):
        j[b * n_block:(b + 1) * n_block, b * n_block:(b + 1) * n_block] = 1
    return j[rng.permutation(n)][:, rng.permutation(n)]


def test_halved_rows_extend_the_exact_range():
    # four blocks of ones J_10 with rows and columns permuted (order 40):
    # perm = (10!)^4 ~ 2^87.  With every y_i = 2 x_i carried whole, T =
    # 2^39 perm and its Bregman bound, 2^126.2, is past the 128-bit range
    # check; every row sum is even, so the walks carry y_i / 2 and
    # T / 2^40 = perm / 2 fits.  Every walk and the generated kernel give
    # the closed form (about 8 s).
    a = _blocks_of_ones()
    want = math.factorial(10) ** 4
    assert O.int_log2_bound(a) >= 126.0  # the oracle's whole-y bound
    assert api.int_exact(a).exact == want
    for walk in ("int32", "fp64", "sparse"):
        assert _walk(a, walk)[0] == want, walk
    assert api.sparse_codegen(a).exact == want
