"""Structural preprocessing (DM pruning, components, DecomposePerman
compressions) on the host, checked without a GPU:

* against the reference's own CPU binary: its "Submatrix count / Average N /
  Average NNZ" line and permanent on the same inputs (tests/golden/
  decomp_golden.json, written by make_decomp_golden.py);
* against the exact oracle: prod over components of the sum of the leaves'
  permanents equals the permanent of the input, on inputs small enough to
  compute whole, with every compression forced (SPM_DECOMP_SMALL)."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR
from oracle import oracle as O
from superman_amd import _lib, api
from superman_amd import workloads as W

DECOMP = json.load(open(os.path.join(GOLDEN_DIR, "decomp_golden.json")))["cases"]


def exact_from_leaves(d):
    if d.structurally_zero:
        return 0
    comp = {}
    for c, m in d.leaves:
        assert np.all(m == np.round(m))
        v, st = O.int_exact(m.astype(np.int64))
        assert st == 0
        comp[c] = comp.get(c, 0) + v
    total = 1
    for v in comp.values():
        total *= v
    return total


@pytest.mark.parametrize("name", sorted(DECOMP), ids=str)
def test_leaves_match_reference(name):
    c = DECOMP[name]
    a = np.load(os.path.join(GOLDEN_DIR, c["file"]), allow_pickle=False)
    d = api.decompose(a)
    assert len(d.leaves) == c["leaves"]
    # printed by the reference with 6 significant digits
    assert float(f"{d.avg_n:.6g}") == c["avg_n"]
    assert float(f"{d.avg_nnz:.6g}") == c["avg_nnz"]


@pytest.mark.parametrize("name", ["c3_sparse01_n40_s40", "c3_sparse01_n40_s42", "c3_sparse01_n40_s44",
                                  "decomp_ints_n38_s102"], ids=str)
def test_leaves_give_reference_permanent(name):
    c = DECOMP[name]
    a = np.load(os.path.join(GOLDEN_DIR, c["file"]), allow_pickle=False)
    v = exact_from_leaves(api.decompose(a))
    # the reference computes in FP64: exact below 2^53 for 0/1 inputs
    assert abs(v - c["value"]) <= 1e-12 * abs(v)
    if c["value"] < 2 ** 53 and name.startswith("c3"):
        assert v == int(c["value"])


def _sparse_case(n, seed, per_row, vmax):
    rng = np.random.default_rng(seed)
    a = np.zeros((n, n))
    p = rng.permutation(n)
    for i in range(n):
        a[i, p[i]] = rng.integers(1, vmax + 1)
        k = int(rng.integers(0, per_row))
        cols = rng.choice([j for j in range(n) if j != p[i]], k, replace=False)
        a[i, cols] = rng.integers(-vmax, vmax + 1, k)
    return a


@pytest.mark.parametrize("seed", range(12))
def test_every_compression_preserves_the_permanent(seed):
    n = 10 + seed % 6
    a = _sparse_case(n, 300 + seed, per_row=4 + seed % 3, vmax=2)
    ref, st = O.int_exact(a.astype(np.int64))
    assert st == 0
    for opts in (_lib.SPM_DECOMP_DEFAULT | _lib.SPM_DECOMP_SMALL, _lib.SPM_DECOMP_COMPRESS | _lib.SPM_DECOMP_SMALL,
                 _lib.SPM_DECOMP_DM | _lib.SPM_DECOMP_COMPONENTS, _lib.SPM_DECOMP_CODEGEN | _lib.SPM_DECOMP_COMPRESS |
                 _lib.SPM_DECOMP_SMALL):
        d = api.decompose(a, opts)
        assert exact_from_leaves(d) == ref, (seed, opts)


def test_small_threshold_really_compresses():
    a = _sparse_case(14, 7, per_row=3, vmax=1)
    d = api.decompose(a, _lib.SPM_DECOMP_DEFAULT | _lib.SPM_DECOMP_SMALL)
    assert max(m.shape[0] for _, m in d.leaves) <= 8 or len(d.leaves) > 1


def test_dm_pruning_removes_entries_outside_perfect_matchings():
    # upper triangular: only the diagonal lies in a perfect matching
    a = np.triu(np.ones((6, 6)))
    d = api.decompose(a, _lib.SPM_DECOMP_DM | _lib.SPM_DECOMP_COMPONENTS)
    assert d.components == 6 and all(m.shape == (1, 1) and m[0, 0] == 1 for _, m in d.leaves)


def test_structurally_zero():
    a = np.ones((5, 5))
    a[:, 2] = 0  # a zero column: no perfect matching
    d = api.decompose(a)
    assert d.structurally_zero and not d.leaves
    a = np.ones((5, 5))
    a[3:, :3] = 0
    a[:3, 3:] = 0
    a[0, :] = 0
    a[0, 4] = 1  # row 0 only reaches the 2x2 corner: unequal component
    assert api.decompose(a).structurally_zero


def test_permutation_invariance():
    a = W.sparse01(40, 40)
    rng = np.random.default_rng(5)
    b = a[rng.permutation(40)][:, rng.permutation(40)]
    assert exact_from_leaves(api.decompose(b)) == 213647004


def test_bad_arguments():
    with pytest.raises(_lib.SpmError):
        api.decompose(np.full((3, 3), np.nan))
    with pytest.raises(_lib.SpmError):
        api.decompose(np.ones((3, 3)), 64)


EXAMPLES = json.load(open(os.path.join(GOLDEN_DIR, "example_golden.json")))["cases"]


@pytest.mark.parametrize("name", sorted(EXAMPLES), ids=str)
def test_example_matrix_leaves_match_reference(name):
    c = EXAMPLES[name]
    a = np.load(os.path.join(GOLDEN_DIR, c["file"]), allow_pickle=False)
    d = api.decompose(a)
    assert len(d.leaves) == c["leaves"]
    assert float(f"{d.avg_n:.6g}") == c["avg_n"]
    assert float(f"{d.avg_nnz:.6g}") == c["avg_nnz"]
    if c["binary"]:  # 0/1: the exact permanent from the leaves, as the reference printed it
        assert exact_from_leaves(d) == int(c["value"])
