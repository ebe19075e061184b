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


@pytest.mark.parametrize("name", [k for k in DECOMP if k.startswith("decomp_ints")], ids=str)
def test_sparse_ints_match_reference(name):
    c = DECOMP[name]
    a = np.load(os.path.join(GOLDEN_DIR, c["file"]), allow_pickle=False)
    r = api.permanent_decomposed(a)
    assert rel_err(r.value, c["value"]) <= 1e-10


def test_dense_real_single_leaf():
    a = W.dense_uniform(20, 8)  # dense: no pruning, one component, no compression
    r = api.permanent_decomposed(a)
    hi, lo = O.dense_dd(a)
    assert rel_err(r.value, hi + lo) <= 1e-12 and r.exact is None


def test_structurally_zero_on_gpu():
    a = np.ones((6, 6))
    a[:, 1] = 0
    assert api.permanent_decomposed(a).exact == 0
