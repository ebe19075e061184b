"""C++ drop-in headers (include/superman): algorithm selection and error
behaviour of CommonFiles/AlgorithmSelector.h:52-139 and
Algorithms/GPU/GPUKernelWrappers.cu, checked without a GPU."""
import numpy as np

import dropin
from superman_amd import workloads as W

XREGISTERMSHARED, XREGISTERMGLOBAL, XGLOBALMGLOBAL = 0, 1, 2
NAIVECG, REGEFFCG, APPROX = 6, 7, 8


def test_auto_selects_dense_for_small_components():
    rc, out = dropin.run("sel", W.dense_uniform(12, 1))
    assert rc == 0 and out[0] == "algorithm %d sparse 0" % XREGISTERMSHARED


def test_auto_follows_the_mi355x_cost_model():
    # the reference's rule would pick code generation here (nnz/row > 6,
    # biggest component 44 > 41 and > n - 3); on MI355X the dense walk is
    # estimated cheaper than compiling and running a 1936-literal kernel
    rc, out = dropin.run("sel", np.ones((44, 44)))
    assert rc == 0 and out[0] == "algorithm %d sparse 0" % XREGISTERMSHARED
    # an order-48 sparse 0/1 matrix: the generated kernel pays for its compile
    rc, out = dropin.run("sel", W.sparse01(48, 48))
    assert rc == 0 and out[0] == "algorithm %d sparse 1" % REGEFFCG


def test_sparse_routing_below_30_percent():
    a = W.sparse01(40, 40)
    assert 100.0 * np.count_nonzero(a) / a.size < 30
    rc, out = dropin.run("sel", a, algorithm=XGLOBALMGLOBAL)
    assert rc == 0 and out[0] == "algorithm %d sparse 1" % XGLOBALMGLOBAL
    rc, out = dropin.run("sel", a, algorithm=XREGISTERMGLOBAL)
    assert rc == 0 and out[0] == "algorithm %d sparse 0" % XREGISTERMGLOBAL


def test_unsupported_algorithm_raises():
    # dense routing with a sparse-only algorithm
    rc, out = dropin.run("run", W.dense_uniform(8, 2), algorithm=XGLOBALMGLOBAL)
    assert rc == 3 and "not included in the available GPU algorithms list" in out[-1]


def test_no_device_is_an_error_not_a_fallback():
    rc, out = dropin.run("run", W.dense_uniform(8, 3), algorithm=XREGISTERMSHARED)
    assert rc == 3 and out[-1].startswith("error spm_permanent_decomposed")
