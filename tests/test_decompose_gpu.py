"""The permanent through the decomposition on the GPU (spm_permanent_decomposed):
exact on the sparse 0/1 n=40 inputs (the reference's values), and equal to
the reference binary's FP64 values on the sparse integer inputs whose leaves
outgrow the exact range."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR, rel_err
from oracle import oracle as O
from superman_amd import api
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu

DECOMP = json.load(open(os.path.join(GOLDEN_DIR, "decomp_golden.json")))["cases"]


@pytest.mark.parametrize("seed", range(40, 45))
def test_sparse01_n40_exact(seed):
    c = DECOMP[f"c3_sparse01_n40_s{seed}"]
    a = np.load(os.path.join(GOLDEN_DIR, c["file"]), allow_pickle=False)
    r = api.permanent_decomposed(a)
    assert r.exact == int(c["value"])


def _truth(a):
    """prod over components of the sum of leaf permanents, each leaf exact
    (int128 oracle) or, beyond its range, double-double."""
    from fractions import Fraction
    comp = {}
    for c, m in api.decompose(a).leaves:
        v, st = O.int_exact(m.astype(np.int64))
        if st != 0:
            hi, lo = O.dense_dd(m)
            v = Fraction(hi) + Fraction(lo)
        comp[c] = comp.get(c, 0) + v
    t = 1
    for v in comp.values():
        t *= v
    return t


@pytest.mark.parametrize("name", [k for k in DECOMP if k.startswith("decomp_ints")], ids=str)
def test_sparse_ints_match_reference(name):
    # leaves whose entries outgrow the exact range run in FP64 (compensated);
    # the bar is 1e-10 against the double-double truth.  The reference's own
    # FP64 value is within 2e-10 of it (1.3e-10 off on decomp_ints_n42_s103).
    c = DECOMP[name]
    a = np.load(os.path.join(GOLDEN_DIR, c["file"]), allow_pickle=False)
    r = api.permanent_decomposed(a)
    t = _truth(a)
    assert abs((r.value - t) / t) <= 1e-10
    assert rel_err(r.value, c["value"]) <= 2e-10


def test_dense_real_single_leaf():
    a = W.dense_uniform(20, 8)  # dense: no pruning, one component, no compression
    r = api.permanent_decomposed(a)
    hi, lo = O.dense_dd(a)
    assert rel_err(r.value, hi + lo) <= 1e-12 and r.exact is None


def test_structurally_zero_on_gpu():
    a = np.ones((6, 6))
    a[:, 1] = 0
    assert api.permanent_decomposed(a).exact == 0
