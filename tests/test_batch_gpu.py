"""Batched persistent kernel and serving loop on the GPU (through the C ABI).

Bars: real and complex within 1e-10 relative of the double-double oracle
(and within the result's own error estimate), integers bit-exact against the
int128 oracle; a request's result is bit-identical alone, in a mixed batch,
eager or graph-launched, fused or per-kind, and served under Poisson
arrivals.  Every engine here (but the one that injects failing checks) ends
with late_publish == 0 and retried == 0: no result arrived late and none was
run again."""
import contextlib

import numpy as np
import pytest

from conftest import rel_err
from oracle import oracle as O
from superman_amd import _lib, api
from superman_amd import workloads as W
from superman_amd.batch import BatchEngine, to_permanent

pytestmark = pytest.mark.gpu

TOL = 1e-10


def make(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "real":
        return rng.random((n, n))
    if kind == "complex":
        return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2.0)
    return (rng.random((n, n)) < 0.5).astype(np.float64)


def check_vs_oracle(kind, a, p):
    if kind == "int":
        ref, st = O.int_exact(a.astype(np.int64))
        assert st == 0
        assert p.exact == ref, (a.shape[0], p.exact, ref)
    elif kind == "real":
        hi, lo = O.dense_dd(a)
        assert rel_err(p.value, hi + lo) <= TOL, (a.shape[0], p.value, hi + lo)
        assert abs(p.value - (hi + lo)) <= p.err_est + 1e-300
    else:
        q = O.complex_dd(a)
        ref = complex(q[0] + q[1], q[2] + q[3])
        assert rel_err(p.value, ref) <= TOL, (a.shape[0], p.value, ref)
        assert abs(p.value - ref) <= p.err_est


@contextlib.contextmanager
def checked(**kw):
    """A BatchEngine whose late-publish and re-run counters must stay 0."""
    with BatchEngine(**kw) as eng:
        yield eng
        assert eng.counters() == {"late_publish": 0, "retried": 0}, eng.last_error()


@pytest.fixture(scope="module")
def engine():
    with checked() as eng:
        yield eng


@pytest.mark.parametrize("kind", ["real", "complex", "int"])
def test_every_order_vs_oracle(engine, kind):
    ns = list(range(1, 23))
    reqs = [(kind, make(kind, n, 100 + n)) for n in ns]
    res = engine.run(reqs)
    for (k, a), p in zip(reqs, res):
        check_vs_oracle(k, a, p)


def test_both_integer_walks_in_one_batch(engine):
    # 0/1 matrices take the FP64 walk, wider entries the int32 walk
    rng = np.random.default_rng(77)
    reqs = []
    for n in range(4, 19):
        reqs.append(("int", rng.integers(-3, 4, (n, n)).astype(np.float64)))
        reqs.append(("int", (rng.random((n, n)) < 0.5).astype(np.float64)))
    res = engine.run(reqs)
    for (k, a), p in zip(reqs, res):
        check_vs_oracle(k, a, p)


def test_large_orders_match_single_matrix_path(engine):
    reqs = []
    for n in range(23, 29):
        reqs += [("real", make("real", n, n)), ("complex", make("complex", n, n)), ("int", make("int", n, n))]
    res = engine.run(reqs)
    for (k, a), p in zip(reqs, res):
        if k == "int":
            assert p.exact == api.int_exact(a.astype(np.int64)).exact
        else:
            # different chunk sizes round differently: both within their
            # error estimates of each other and within the 1e-10 bar
            s = api.dense_f64(a) if k == "real" else api.dense_c128(a)
            assert abs(p.value - s.value) <= p.err_est + s.err_est
            assert rel_err(p.value, s.value) <= TOL


def test_closed_forms(engine):
    import math
    reqs = [("real", np.ones((n, n))) for n in (5, 12, 20)]
    reqs += [("complex", 1j * np.ones((n, n))) for n in (6, 11)]
    reqs += [("int", np.ones((n, n))) for n in (9, 18, 24)]
    res = engine.run(reqs)
    for (k, a), p in zip(reqs, res):
        n = a.shape[0]
        if k == "int":
            assert p.exact == math.factorial(n)
        elif k == "real":
            assert rel_err(p.value, math.factorial(n)) <= 1e-14
        else:
            assert rel_err(p.value, (1j) ** n * math.factorial(n)) <= 1e-14


def _mix(count, seed):
    s = W.serving_stream(count, rate=1.0, seed=seed, n_lo=8, n_hi=24)
    return [(q.kind, q.a, q.rid) for q in s]


def _same(p, q):
    if p.exact is not None or q.exact is not None:
        return p.exact == q.exact
    return p.hi == q.hi and p.lo == q.lo


def test_alone_equals_in_mix(engine):
    mix = _mix(120, 11)
    together = engine.run(mix)
    for i in range(0, 120, 7):
        alone = engine.run([mix[i]])[0]
        assert _same(alone, together[i]), (i, alone.value, together[i].value)


def test_graph_replays_equal_eager():
    mix = _mix(60, 12)
    with checked(use_graph=False) as eng:
        eager = eng.run(mix)
    with checked(use_graph=True) as eng:
        for rep in range(3):
            got = eng.run(mix)
            assert all(_same(a, b) for a, b in zip(got, eager)), rep
            # eager work between replays
            api.dense_f64(make("real", 12, rep))


def test_fused_equals_per_kind():
    mix = _mix(60, 13)
    with checked() as eng:
        t = eng.tuning
        t.fused = False
        eng.tuning = t
        split = eng.run(mix)
        t.fused = True
        eng.tuning = t
        fused = eng.run(mix)
    for a, b in zip(split, fused):
        assert _same(a, b) or rel_err(a.value, b.value) <= 1e-15


def test_chunk_size_changes_rounding_only():
    mix = _mix(40, 14)
    with checked() as eng:
        r1 = eng.run(mix)
        t = eng.tuning
        t.log2_chunk = {"real": 7, "complex": 6, "int": 8}
        eng.tuning = t
        r2 = eng.run(mix)
    for (k, a, _), p, q in zip(mix, r1, r2):
        if k == "int01":
            assert p.exact == q.exact
        else:
            assert abs(p.value - q.value) <= p.err_est + q.err_est


def test_small_batches_and_slots():
    mix = _mix(50, 15)
    with checked(slots=3, max_requests=4) as eng:
        got = eng.run(mix)
    with checked(slots=1) as eng:
        ref = eng.run(mix)
    assert all(_same(a, b) for a, b in zip(got, ref))


def test_serving_poisson_results_equal_alone():
    stream = W.serving_stream(300, rate=3000.0, seed=16, n_lo=8, n_hi=22)
    reqs = [(q.kind, q.a, q.rid) for q in stream]
    arr = np.array([q.arrival for q in stream])
    with checked(slots=2) as eng:
        rc, res, lat, stats = eng.replay(reqs, arr)
        assert rc == 0 and stats.completed == 300
        assert stats.batches > 1  # requests arrived while others ran
        for i in range(0, 300, 13):
            alone = eng.run([reqs[i]])[0]
            assert _same(alone, to_permanent(res[i], reqs[i][0])), i
    assert (lat > 0).all() and stats.p99_s >= stats.p50_s


def test_submit_poll_cancel_on_device():
    with checked(slots=1, max_requests=1) as eng:
        big = make("complex", 26, 1)
        small = make("real", 10, 2)
        t1 = eng.submit("complex", big)
        t2 = eng.submit("real", small)
        t3 = eng.submit("real", small)
        eng.progress()
        assert eng.cancel(t2) == 0
        got = {}
        import time
        t0 = time.time()
        while len(got) < 3 and time.time() - t0 < 60:
            for tk, tag, st, p, c in eng.poll():
                got[tk] = (st, p)
        assert got[t2][0] == _lib.SPM_ECANCELLED
        assert got[t1][0] == 0 and got[t3][0] == 0
        check_vs_oracle("real", small, got[t3][1])


def test_autotune_keeps_results(engine):
    sample = _mix(80, 17)
    before = engine.run(sample)
    with checked() as eng:
        best, sec = eng.autotune(sample, reps=1)
        assert sec > 0 and 2 <= min(best.log2_chunk.values())
        after = eng.run(sample)
    for (k, a, _), p, q in zip(sample, before, after):
        if k == "int01":
            assert p.exact == q.exact
        else:
            assert abs(p.value - q.value) <= p.err_est + q.err_est


def test_rerun_after_a_failed_check_gives_the_same_results():
    stream = W.serving_stream(60, rate=1.0, seed=21, n_lo=8, n_hi=20)
    reqs = [(q.kind, q.a, q.rid) for q in stream]
    with BatchEngine(slots=2, max_requests=8) as eng:
        rc, first, _, st = eng.replay(reqs, np.zeros(60))
        assert rc == 0
        eng.inject_bad_results(5)
        rc, again, _, st = eng.replay(reqs, np.zeros(60))
        assert rc == 0 and st.retried == 5 and st.completed == 60 and st.late_publish == 0
        assert eng.counters() == {"late_publish": 0, "retried": 5}
        for q, a, b in zip(reqs, first, again):
            assert _same(to_permanent(a, q[0]), to_permanent(b, q[0]))


def test_many_batches_every_result_checked():
    # hundreds of batches over two slots, every result against the one-call
    # path: integers equal, floating point within the two error estimates
    stream = W.serving_stream(3000, rate=4e4, seed=33, n_lo=8, n_hi=22)
    reqs = [(q.kind, q.a, q.rid) for q in stream]
    arr = np.array([q.arrival for q in stream])
    with checked(slots=2, max_requests=32) as eng:
        rc, res, _, st = eng.replay(reqs, arr)
    assert rc == 0 and st.completed == 3000 and st.batches >= 3000 // 32  # at most 32 per batch
    calls = {"real": api.dense_f64, "complex": api.dense_c128, "int01": lambda a: api.int_exact(a.astype(np.int64))}
    for (kind, a, _), r in zip(reqs, res):
        b = to_permanent(r, kind)
        p = calls[kind](a)
        if kind == "int01":
            assert b.exact == p.exact
        else:
            assert abs(b.value - p.value) <= b.err_est + p.err_est


def test_serving_poisson_full_order_range():
    # the serving mix's whole order range, n = 8..28: every result against
    # the one-call path (integers equal, floating point within both error
    # estimates and the 1e-10 bar)
    stream = W.serving_stream(400, rate=8000.0, seed=41, n_lo=8, n_hi=28)
    reqs = [(q.kind, q.a, q.rid) for q in stream]
    arr = np.array([q.arrival for q in stream])
    with checked(slots=2) as eng:
        rc, res, lat, st = eng.replay(reqs, arr)
    assert rc == 0 and st.completed == 400 and st.batches > 1
    assert max(a.shape[0] for _, a, _ in reqs) == 28
    calls = {"real": api.dense_f64, "complex": api.dense_c128, "int01": lambda a: api.int_exact(a.astype(np.int64))}
    for (kind, a, _), r in zip(reqs, res):
        b = to_permanent(r, kind)
        p = calls[kind](a)
        if kind == "int01":
            assert b.exact == p.exact
        else:
            assert abs(b.value - p.value) <= b.err_est + p.err_est
            assert rel_err(b.value, p.value) <= TOL
