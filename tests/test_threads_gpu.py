"""The C ABI called from several host threads at once: every entry point is
reentrant (own buffers, streams and events per call; the generated-kernel
cache and device table behind locks), so concurrent results equal the
sequential ones bit for bit."""
import concurrent.futures as cf

import numpy as np
import pytest

from superman_amd import api
from superman_amd import workloads as W

pytestmark = pytest.mark.gpu


def _jobs():
    jobs = []
    for i in range(6):
        jobs.append(("real", W.dense_normal(18 + i, 50 + i)))
        jobs.append(("complex", W.complex_gauss(14 + i, 60 + i)))
        jobs.append(("int", W.sparse01(26 + i, 70 + i).astype(np.int64)))
        jobs.append(("decomposed", W.sparse01(36 + i, 80 + i)))
    jobs.append(("codegen", W.sparse01(24, 5).astype(np.int64)))
    jobs.append(("codegen", W.sparse01(24, 5).astype(np.int64)))  # same pattern: one cache entry
    return jobs


def _run(job):
    kind, a = job
    if kind == "real":
        r = api.dense_f64(a)
        return (r.hi, r.lo)
    if kind == "complex":
        r = api.dense_c128(a)
        return (r.hi, r.lo)
    if kind == "int":
        return api.int_exact(a).exact
    if kind == "decomposed":
        r = api.permanent_decomposed(a)
        return r.exact if r.exact is not None else r.value
    return api.sparse_codegen(a).exact


def test_concurrent_calls_equal_sequential():
    jobs = _jobs()
    seq = [_run(j) for j in jobs]
    with cf.ThreadPoolExecutor(max_workers=6) as ex:
        for _ in range(2):
            par = list(ex.map(_run, jobs))
            assert par == seq
