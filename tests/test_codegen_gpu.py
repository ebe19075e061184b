"""Generated SpaRyser kernels (hiprtc) on the GPU: bit-exact against the
int128 oracle, the reference's own outputs and the generic integer kernel."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, golden_matrix, load_golden
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


def two_per_row(n, seed, signed):
    """A seeded permutation plus one more entry per row; +-1 entries when
    signed.  Half of a row's subset sums are zero, so a term is nonzero in
    about 2^-n of the subsets."""
    rng = np.random.default_rng(seed)
    a = np.zeros((n, n), dtype=np.int64)
    p = rng.permutation(n)
    for i in range(n):
        a[i, p[i]] = 1
        a[i, rng.choice([j for j in range(n) if j != p[i]])] = 1
    if signed:
        a *= rng.choice([-1, 1], size=(n, n))
    return a


@pytest.mark.parametrize("n,seed,signed", [(24, 324, False), (24, 325, True), (26, 326, False), (26, 326, True),
                                           (30, 330, False), (30, 334, True)])
def test_rarest_terms_tested_per_block(n, seed, signed):
    # patterns whose estimated nonzero-term fraction is below 5e-7 test their
    # terms once per block and form them on a walk back through the block
    # (8-step blocks at these orders, 32 from order 31): exact against the
    # int128 oracle, signs included
    a = two_per_row(n, seed, signed)
    assert "Zm = min(Zm, Z);" in api.codegen_source(a)
    ref, st = O.int_exact(a)
    assert st == 0 and ref != 0
    assert api.sparse_codegen(a).exact == ref


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


def test_disk_cache_across_processes(tmp_path):
    # a generated kernel compiled in one process is loaded from the on-disk
    # cache by the next one: same result, a load instead of a compile.
    # hiprtc's own cache (comgr's) is switched off, so that the first
    # process compiles even when an earlier test built the same source.
    code = ("import json, sys; sys.path.insert(0, %r); import numpy as np; "
            "from superman_amd import api, workloads as W; a = W.sparse01(30, 7).astype(np.int64); "
            "p = api.CodegenPlan(a); r = p.run(); i = p.codegen_info(); p.close(); "
            "print(json.dumps({'exact': str(r.exact), 'compile_seconds': i['compile_seconds']}))") % ROOT
    env = dict(os.environ, SPM_KERNEL_CACHE=str(tmp_path), AMD_COMGR_CACHE="0")
    runs = []
    for _ in range(2):
        p = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stderr
        runs.append(json.loads(p.stdout.strip().splitlines()[-1]))
    assert len(list(tmp_path.glob("spm_gen_*.co"))) == 1
    assert runs[0]["exact"] == runs[1]["exact"]
    assert runs[1]["compile_seconds"] < 0.2 * runs[0]["compile_seconds"]


def test_auto_choice_counts_a_compiled_kernel_as_a_load(tmp_path, monkeypatch):
    # order 40, 4 entries per row: a compile (0.3 s) does not pay for one
    # permanent, but once the pattern's kernel is compiled the automatic
    # path takes it (about 0.05 s against 0.23 s for the compressed-column
    # walk);
    # spm_choose_path, the choice for kernels not yet compiled, stays put.
    # An empty on-disk cache of its own, and a pattern no other test uses.
    monkeypatch.setenv("SPM_KERNEL_CACHE", str(tmp_path))
    a = W.sparse01(40, 4040).astype(np.float64)
    r0, path0 = api.permanent_auto(a)
    assert path0 == "int"
    with api.CodegenPlan(a.astype(np.int64)) as p:
        assert p.run().exact == r0.exact
    r1, path1 = api.permanent_auto(a)
    assert path1 == "gen_int" and r1.exact == r0.exact
    assert api.choose_path(a) == "int"
