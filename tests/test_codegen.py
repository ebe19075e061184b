"""SpaRyser code generation checks that need no GPU (CPU only)."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR
from superman_amd import api
from superman_amd import workloads as W


def test_ut_order_matches_reference_generated_kernel():
    # The reference's checked-in generated kernel was built from bcsstk01
    # after IO::UTOrder; every column's rows and (6-decimal) values must match.
    a = np.load(os.path.join(GOLDEN_DIR, "bcsstk01_undirected.npy"), allow_pickle=False)
    ref = json.load(open(os.path.join(GOLDEN_DIR, "ut_bcsstk01_reference.json")))["columns"]
    rp, cp = api.ut_order(a)
    b = api.apply_permutations(a, rp, cp)
    n = a.shape[0]
    assert len(ref) == n
    for j in range(n):
        mine = sorted([i, "%f" % b[i, j]] for i in range(n) if b[i, j] != 0)
        assert mine == ref[str(j)], j


def test_ut_order_is_sparsest_first_permutation():
    a = W.sparse01(30, 5)
    rp, cp = api.ut_order(a)
    assert sorted(rp) == list(range(30)) and sorted(cp) == list(range(30))
    b = api.apply_permutations(a, rp, cp)
    # column 0 of the ordered matrix is a column of least in-degree
    assert (b[:, 0] != 0).sum() == min((a[:, j] != 0).sum() for j in range(30))
    # rows are numbered in first-touch order: row r's first nonzero column is
    # non-decreasing in r
    first = [int(np.flatnonzero(b[i])[0]) for i in range(30)]
    assert first == sorted(first)


def test_generated_literals_are_exact():
    a = np.eye(6, dtype=np.int64) * 123456789 + 1
    src = api.codegen_source(a, register_efficient=False)
    assert "+= 246913580;" in src or "+= 246913580" in src  # 2 * (123456789 + 1)
    assert ".0000" not in src  # no rounded decimal literals


def test_generated_source_golden_5x5():
    a = np.array([[1, 1, 0, 0, 1], [0, 1, 1, 0, 0], [1, 0, 1, 1, 0], [0, 1, 0, 1, 1], [1, 0, 1, 0, 1]])
    src = api.codegen_source(a)
    golden = os.path.join(GOLDEN_DIR, "codegen_5x5.txt")
    assert src == open(golden).read()


@pytest.mark.parametrize("n,seed", [(5, 1), (12, 2), (40, 40)])
def test_generated_source_compiles(n, seed):
    a = W.sparse01(n, seed) if n > 5 else np.ones((5, 5), dtype=np.int64)
    info = api.codegen_compile_only(a)
    assert info["code_bytes"] > 1000


def rows_of(n, per_row, seed):
    """0/1, a seeded permutation plus per_row - 1 further entries per row."""
    rng = np.random.default_rng(seed)
    a = np.zeros((n, n), dtype=np.int64)
    perm = rng.permutation(n)
    for i in range(n):
        a[i, perm[i]] = 1
        others = rng.choice([j for j in range(n) if j != perm[i]], size=per_row - 1, replace=False)
        a[i, others] = 1
    return a


@pytest.mark.parametrize("n,per_row,per_block,waves", [(40, 4, True, 5), (40, 6, True, 5), (40, 8, False, 5),
                                                       (36, 6, False, 5), (44, 6, True, 5), (44, 8, False, 4)])
def test_term_test_per_block_or_per_step(n, per_row, per_block, waves):
    # the estimated nonzero-term fraction (row-independence estimate; rows of
    # 4 / 6 / 8 entries at order 40: 6.8e-9 / 3.1e-7 / 2.8e-6) picks one term
    # test per 32-step block below 5e-7, else one per step; kernels testing
    # per block ask for 5 waves per SIMD up to order 48, the others for 4
    # past order 40
    src = api.codegen_source(rows_of(n, per_row, 7000 * n + per_row))
    assert ("Zm = min(Zm, Z);" in src) == per_block
    assert ("ballot_w64(Zm == 0)" in src) == per_block
    assert "amdgpu_waves_per_eu(%d)" % waves in src


def test_codegen_rejects_non_integers():
    with pytest.raises(ValueError):
        api.codegen_source(np.full((6, 6), 0.5))


def _real_sparse(n, seed, per_row=4):
    rng = np.random.default_rng(seed)
    a = W.sparse01(n, seed) * rng.standard_normal((n, n))
    a[a == 0] = 0.0
    return a


def test_real_source_literals_are_exact():
    # every hex-float literal of the generated source parses back to one of
    # the matrix's entries (or its negation), bit for bit, and every entry
    # appears: nothing is rounded to a decimal string
    import re
    a = _real_sparse(12, 5)
    a[3, 4] = 0.1  # not dyadic: a 6-decimal literal would be exact only by luck
    a[7, 1] = -1.0 / 3.0
    src = api.codegen_source_real(a, register_efficient=False)  # input order: column n-1 is never flipped
    lits = {float.fromhex(t) for t in re.findall(r"-?0x[0-9a-f]+(?:\.[0-9a-f]*)?p[+-]\d+", src)}
    vals = {float(v) for v in a[:, :-1].ravel() if v != 0}
    assert vals <= {abs(v) for v in lits} | {-abs(v) for v in lits}
    flips = set(re.findall(r"x\d+ \+= (-?0x[0-9a-f.]+p[+-]\d+);", src))
    for t in flips:
        v = float.fromhex(t)
        assert v in vals or -v in vals, t
    assert "0.1" not in src.replace("0x1", "")  # no decimal spelling of the entries


def test_real_source_levels_follow_ut_staircase():
    # a banded pattern in UT order: column c touches rows <= c + 2, so the
    # generator keeps more than one suffix-product level
    n = 24
    a = np.zeros((n, n))
    rng = np.random.default_rng(3)
    for i in range(n):
        for j in range(max(0, i - 2), min(n, i + 3)):
            a[i, j] = rng.standard_normal()
    src = api.codegen_source_real(a)
    m = __import__("re").search(r"(\d+) suffix-product levels", src)
    assert m and int(m.group(1)) >= 3


def test_real_source_compiles_for_gfx950(tmp_path):
    import subprocess
    a = _real_sparse(20, 9)
    src = api.codegen_source_real(a, precision="kahan")
    f = tmp_path / "gen.hip"
    f.write_text("#include <hip/hip_runtime.h>\n" + src)
    out = subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O1", "-std=c++17", "--cuda-device-only",
                          "-c", str(f), "-o", str(tmp_path / "gen.o")], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr[-2000:]
