"""Orders past the pipelined dense walk, the reference's closed-form example
matrices, and accuracy evidence at order 40 that shares no products with
the default run, on the GPU (through the C ABI and the command line).

Bars: closed forms within each result's error estimate, and within 1e-13
relative where every term has the same magnitude (diagonal matrices);
invariances between independent walks (transpose, row and column
permutations, another chunk length) within the two results' error
estimates, and 1e-10 relative where the estimates are below it (not for the
complex Gaussian order-37 case)."""
import json
import math
import os

import numpy as np
import pytest

import cli
from conftest import GOLDEN_DIR, golden_matrix, rel_err
from superman_amd import api
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu


def closed(name):
    with open(os.path.join(GOLDEN_DIR, "closed_form_examples.json")) as f:
        return json.load(f)[name]


def example(name):
    return np.load(os.path.join(GOLDEN_DIR, f"example_{name}.npy"), allow_pickle=False)


def test_n40_other_chunk_length(golden):
    # chunks of 2^(L-6) steps rebuild x exactly at 64x more places: every
    # term's factors carry other roundings than in the default run
    a = golden_matrix(golden["c2_dense_normal_n40_s40"])
    with api.DensePlan(a, "kahan") as p:
        r0 = p.run()
        L0 = p.info()["log2_chunk"]
        p.set_log2_chunk(L0 - 6)
        assert p.info()["log2_chunk"] == L0 - 6
        r1 = p.run()
    assert abs(r0.value - r1.value) <= r0.err_est + r1.err_est
    assert rel_err(r1.value, r0.value) <= 1e-10


def test_f32_product_n40_uniform(golden):
    # the FP32-product mode on the nonnegative n=40 input against the
    # compensated FP64 walk: within the sum of the two stated estimates
    u = golden_matrix(golden["c2_dense_uniform_n40_s40"])
    with api.DensePlan(u, "kahan") as p:
        r64 = p.run()
    with api.DensePlan(u, "f32_product") as p:
        rf = p.run()
    assert abs(rf.value - r64.value) <= rf.err_est + r64.err_est
    assert rel_err(rf.value, r64.value) <= 1e-2


def test_order_41_diagonal():
    # x_i(S) = +-d_i / 2 for every subset: every term is +-prod(d) / 2^n
    d = W.dense_normal(41, 4141).diagonal().copy()
    r = api.dense_f64(np.diag(d), "kahan")
    assert rel_err(r.value, np.prod(d)) <= 1e-13


def test_order_41_invariances():
    # the unpipelined walk with 32-value column heads and tail loads
    a = W.dense_normal(41, 41)
    rng = np.random.default_rng(41)
    r = api.dense_f64(a, "kahan")
    rt = api.dense_f64(np.ascontiguousarray(a.T), "kahan")
    rp = api.dense_f64(np.ascontiguousarray(a[rng.permutation(41)][:, rng.permutation(41)]), "kahan")
    for s in (rt, rp):
        assert abs(s.value - r.value) <= r.err_est + s.err_est
        assert rel_err(s.value, r.value) <= 1e-10


def test_complex_order_37():
    d = W.complex_gauss(37, 3737).diagonal().copy()
    r = api.dense_c128(np.diag(d))
    assert rel_err(r.value, np.prod(d)) <= 1e-13
    a = W.complex_gauss(37, 370)
    r1, r2 = api.dense_c128(a), api.dense_c128(np.conj(a))
    assert abs(r2.value - np.conj(r1.value)) <= r1.err_est + r2.err_est
    rt = api.dense_c128(np.ascontiguousarray(a.T))
    # a complex Gaussian order-37 sum cancels enough that the estimates
    # themselves pass 1e-10 relative (transposes measured 2.2e-10 apart)
    assert abs(rt.value - r1.value) <= r1.err_est + rt.err_est


def test_example_complex_mtx():
    # ExampleMatrices/complex.mtx: 40 x 40, every entry i; perm = 40! i^40
    a = example("complex")
    ref = float(closed("complex")["perm"])
    r = api.dense_c128(a)
    assert abs(r.value - ref) <= r.err_est
    rc, out, err = cli.run(a, "complex=true")
    assert rc == 0, err
    assert cli.value(out) == r.value  # the command line runs the same kernel


def test_example_dense_mtx():
    # ExampleMatrices/dense.mtx: 45 x 45, every entry 0.91; perm = 45! 0.91^45
    # (with 0.91 the binary64 value).  About 45 s on the dense walk.  Every
    # x_i(S) is 0.91 (k - 21.5) for |S| = k, so sum |term| has a closed form,
    # and with it the compensated result's error estimate 2 n u sum|term|
    # (relative 1.1e-7: the all-equal matrix cancels heavily).
    a = example("dense")
    ref = closed("dense")["perm"]
    n = a.shape[0]
    from fractions import Fraction
    abs_sum = 2 * sum(math.comb(n - 1, k) * abs(Fraction(2 * k - (n - 2), 2) * Fraction(0.91)) ** n
                      for k in range(n))
    est = 2 * n * 2.0 ** -53 * float(abs_sum)
    rc, out, err = cli.run(a)
    assert rc == 0, err
    v = cli.value(out)
    assert abs(v - ref) <= est, (v, ref, est)
