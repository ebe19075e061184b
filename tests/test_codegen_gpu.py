"""Generated SpaRyser kernels (hiprtc) on the GPU: bit-exact against the
int128 oracle, the reference's own outputs and the generic integer kernel."""
import numpy as np
import pytest

from conftest import golden_matrix, load_golden
from oracle import oracle as O
from superman_amd import api
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("register_efficient", [True, False])
@pytest.mark.parametrize("case", [c for c in load_golden() if c["name"].startswith(("sparse01_n", "int01_d50"))
                                  and c["n"] >= 5], ids=lambda c: c["name"])
def test_golden_vs_oracle(case, register_efficient):
    a = golden_matrix(case)
    r = api.sparse_codegen(a, register_efficient)
    assert r.exact == int(case["exact"])


@pytest.mark.parametrize("case", [c for c in load_golden() if c["name"].startswith("c3_sparse01_n40")],
                         ids=lambda c: c["name"])
def test_sparse01_n40_matches_reference(case):
    r = api.sparse_codegen(golden_matrix(case))
    assert r.exact == int(case["ref8"])


@pytest.mark.parametrize("n", [5, 6, 9, 14, 19])
def test_signed_sparse_ints(n):
    rng = np.random.default_rng(700 + n)
    a = rng.integers(-3, 4, (n, n)) * (rng.random((n, n)) < 0.4)
    a[np.arange(n), rng.permutation(n)] = rng.integers(1, 4, n)
    ref, st = O.int_exact(a)
    assert st == 0
    assert api.sparse_codegen(a).exact == ref
    assert api.int_exact(a).exact == ref


def test_cache_reuses_compiled_kernel():
    a = W.sparse01(24, 99)
    with api.CodegenPlan(a) as p1:
        r1 = p1.run()
        i1 = p1.codegen_info()
    with api.CodegenPlan(a) as p2:
        r2 = p2.run()
        i2 = p2.codegen_info()
    assert r1.exact == r2.exact == O.int_exact(a.astype(np.int64))[0]
    assert i2["cached"] and i2["hash"] == i1["hash"]


def test_order_42_matches_sparse_walk():
    # orders 41..48 ask for 4 waves per SIMD: the generated kernel and the
    # compressed-column walk (independent kernels) agree exactly
    a = W.sparse01(42, 42).astype(np.int64)
    with api.CodegenPlan(a) as p:
        g = p.run().exact
    with api.DensePlan(a, "exact", int_walk="sparse") as q:
        s = q.run().exact
    assert g == s and g > 0


def test_order_44_triangular_closed_form():
    # a generated kernel past order 42 against a closed form: an upper-
    # triangular 0/1 pattern (ones on the diagonal, 3 more entries per row
    # above it) has perm = 1, so the 128-bit sum of 2^43 terms must cancel
    # to exactly 2^43 (about 4 s with the compile)
    n = 44
    rng = np.random.default_rng(n)
    a = np.zeros((n, n), dtype=np.int64)
    for i in range(n):
        a[i, i] = 1
        above = np.arange(i + 1, n)
        if above.size:
            a[i, rng.choice(above, size=min(3, above.size), replace=False)] = 1
    assert api.sparse_codegen(a).exact == 1
