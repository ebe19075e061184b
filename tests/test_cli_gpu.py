"""The command-line driver on the GPU: main.cpp's flow (DM pruning,
components, algorithm selection, per-component product) end to end."""
import json
import math
import os

import numpy as np
import pytest

import cli
from conftest import GOLDEN_DIR, rel_err
from oracle import oracle as O
import superman_amd as S
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("algorithm", ["auto", "xregister_mshared", "xglobal_mglobal",
                                       "register_efficient_code_generation", "naive_code_generation"])
def test_sparse01_n40(algorithm):
    a = np.load(os.path.join(GOLDEN_DIR, "c3_sparse01_n40_s40.npy"), allow_pickle=False)
    rc, out, err = cli.run(a, f"algorithm={algorithm}")
    assert rc == 0, err
    assert cli.value(out) == 213647004.0
    if algorithm == "auto":  # the reference binary's line for this input (5 leaves)
        assert "Submatrix count: 5 - Average N: 18.4 - Average NNZ: 78.6" in out


def test_pattern_file_reads_ones():
    a = np.ones((10, 10))
    rc, out, err = cli.run(a, pattern=True)
    assert rc == 0, err
    assert cli.value(out) == float(math.factorial(10))


@pytest.mark.parametrize("precision", ["dd", "kahan"])
def test_dense_real(precision):
    a = W.dense_uniform(14, 3)
    rc, out, err = cli.run(a, f"calculation_precision={precision}")
    assert rc == 0, err
    hi, lo = O.dense_dd(a)
    assert rel_err(cli.value(out), hi + lo) <= 1e-12


def test_complex():
    a = 1j * np.ones((12, 12))
    rc, out, err = cli.run(a, "complex=true")
    assert rc == 0, err
    assert rel_err(cli.value(out), (1j) ** 12 * math.factorial(12)) <= 1e-14


def test_undirected_mirrors_entries():
    a = np.tril(np.ones((8, 8)))  # undirected read -> all ones
    rc, out, err = cli.run(a, "undirected=true")
    assert rc == 0, err
    assert cli.value(out) == float(math.factorial(8))


EXAMPLES = json.load(open(os.path.join(GOLDEN_DIR, "example_golden.json")))["cases"]


@pytest.mark.parametrize("name", sorted(EXAMPLES))
def test_reference_example_matrices(name):
    # ExampleMatrices/<name>.mtx as its banner describes it; the reference's
    # own CPU binary's permanent on the same input
    c = EXAMPLES[name]
    a = np.load(os.path.join(GOLDEN_DIR, c["file"]), allow_pickle=False)
    args = [f"binary={'true' if c['binary'] else 'false'}", f"undirected={'true' if c['undirected'] else 'false'}"]
    rc, out, err = cli.run(a, *args, pattern=c["binary"], symmetric=c["undirected"])
    assert rc == 0, err
    v = cli.value(out)
    if c["binary"]:  # integer result below 2^53: exact in both
        assert v == c["value"]
    else:
        # d_ss is ill-conditioned (entries from 1e-5 to 2.5e4): FP64 results
        # carry errors of order cond * n * 2^-53, the reference's too, so
        # the check is against the leaves' double-double permanent within
        # the library's own bound for this input
        r = S.permanent_decomposed(a)
        truth = c["leaves_dd"]
        print(f"{name}: cli {v!r} library {r.value!r} dd {truth!r} reference {c['value']!r} "
              f"err_est/|dd| {r.err_est / abs(truth):.2e}")
        assert abs(r.value - truth) <= r.err_est
        assert abs(v - truth) <= r.err_est
        # compensated sums: closer to the truth than the reference's FP64 value
        assert abs(v - truth) < abs(c["value"] - truth)


def _scaling_cases():
    with open(os.path.join(GOLDEN_DIR, "scaling_golden.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _scaling_cases(), ids=lambda c: c["name"])
@pytest.mark.parametrize("algorithm", ["auto", "register_efficient_code_generation"])
def test_scaling_matches_reference(case, algorithm):
    # scaling=true: every leaf balanced (IO::scale) before its permanent and
    # the scales divided out after it; the reference binary's printed value
    # and the unscaled truth agree with ours within 1e-10 relative
    a = getattr(W, case["generator"])(case["n"], case["seed"])
    rc, out, err = cli.run(a, "scaling=true", f"scaling_iteration_no={case['scaling_iteration_no']}",
                           f"scale_into={case['scale_into']}", f"algorithm={algorithm}")
    assert rc == 0, err
    v = cli.value(out)
    if case["generator"] == "sparse01":  # an integer permanent: the reference's unscaled value is exact
        truth = case["reference_plain"]
    else:
        hi, lo = O.dense_dd(a)
        truth = hi + lo
    assert rel_err(v, truth) <= 1e-10, (v, truth)
    assert rel_err(v, case["reference_scaled"]) <= 1e-10, (v, case["reference_scaled"])


def test_scaled_decomposed_api_vs_unscaled():
    a = W.sparse01(36, 36) * (1.0 + W.dense_uniform(36, 37))
    r0 = S.permanent_decomposed(a)
    r1 = S.permanent_decomposed(a, scaling=(100, 2.0))
    assert r1.exact is None
    assert abs(r1.value - r0.value) <= 2 * (r0.err_est + r1.err_est) + 1e-10 * abs(r0.value)
