"""C++ drop-in headers on the GPU: a program written against the
reference's interfaces (AlgorithmSelector -> gpuDPSingleGPU /
gpuSPSingleGPU, gpuComputeComplex) returns the oracle's permanents."""
import numpy as np
import pytest

import dropin
from conftest import golden_matrix, rel_err
from oracle import oracle as O
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("algorithm", [0, 1, 4, 5])
def test_dense_algorithms(algorithm, precision):
    a = W.dense_uniform(18, 5)
    rc, out = dropin.run("run", a, algorithm=algorithm, precision=precision)
    assert rc == 0, out
    hi, lo = O.dense_dd(a)
    assert rel_err(float(out[1].split()[0]), hi + lo) < 1e-12


@pytest.mark.parametrize("algorithm", [2, 3, 4, 5, 6, 7])
def test_sparse_algorithms_exact(algorithm, golden):
    c = golden["sparse01_n24_s24"]
    a = golden_matrix(c)
    rc, out = dropin.run("run", a, algorithm=algorithm)
    assert rc == 0, out
    assert float(out[1].split()[0]) == float(int(c["exact"]))


def test_auto_selection_runs():
    a = W.dense_uniform(20, 7)
    rc, out = dropin.run("run", a)
    assert rc == 0, out
    assert out[0] == "algorithm 0 sparse 0"
    hi, lo = O.dense_dd(a)
    assert rel_err(float(out[1].split()[0]), hi + lo) < 1e-12


def test_complex():
    a = W.complex_gauss(14, 3)
    rc, out = dropin.run("cplx", a)
    assert rc == 0, out
    re, im = map(float, out[0].split())
    ref = O.complex_dd(a)
    assert abs(complex(re, im) - complex(ref[0] + ref[1], ref[2] + ref[3])) <= 1e-11 * abs(
        complex(ref[0] + ref[1], ref[2] + ref[3]))
