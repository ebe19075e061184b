"""Serving-loop logic without a GPU: the engine's dry-run executor forms and
"runs" batches (for a time proportional to their work) without any device
call and completes them with NaN values, so admission, budgets, deadlines,
cancellation and fault handling can be tested here."""
import time

import numpy as np
import pytest

from superman_amd import _lib
from superman_amd import workloads as W
from superman_amd.batch import BatchEngine


def reqs(count, seed=0, n_lo=8, n_hi=20):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(n_lo, n_hi + 1))
        kind = ("real", "complex", "int")[i % 3]
        a = rng.random((n, n)) if kind != "int" else (rng.random((n, n)) < 0.5).astype(float)
        if kind == "complex":
            a = a + 1j * rng.random((n, n))
        out.append((kind, a, i))
    return out


def drain(eng, timeout=10.0):
    got = []
    t = time.time()
    while time.time() - t < timeout:
        got += eng.poll()
        q, f = eng.pending()
        if q == 0 and f == 0:
            got += eng.poll()
            return got
        time.sleep(1e-4)
    raise AssertionError("engine did not drain")


def test_run_completes_every_request():
    with BatchEngine(dry_run=True) as eng:
        st, res = eng.run_raw(reqs(50))
        assert st == 0
        assert all(r.status == 0 for r in res)
        assert all(np.isnan(r.re) for r in res)  # dry run computes nothing


def test_queue_full_is_refused():
    with BatchEngine(dry_run=True, max_pending=4, slots=1) as eng:
        a = np.ones((10, 10))
        sts = [eng.try_submit("real", a)[0] for _ in range(6)]
        # nothing has been launched yet (no progress call): 4 queued, 2 refused
        assert sts[:4] == [0] * 4 and sts[4:] == [_lib.SPM_EAGAIN] * 2
        eng.progress()  # launches the queued ones
        assert eng.try_submit("real", a)[0] == 0
        drain(eng)


def test_never_fits_is_enomem():
    with BatchEngine(dry_run=True, arena_bytes=4096) as eng:
        st, _ = eng.try_submit("real", np.ones((28, 28)))
        assert st == _lib.SPM_ENOMEM
        assert eng.try_submit("real", np.ones((8, 8)))[0] == 0
        drain(eng)


def test_bad_requests():
    with BatchEngine(dry_run=True) as eng:
        assert eng.try_submit("real", np.ones((29, 29)))[0] == _lib.SPM_EINVAL
        bad = np.ones((5, 5))
        bad[2, 2] = np.inf
        assert eng.try_submit("real", bad)[0] == _lib.SPM_EINVAL
        assert eng.try_submit("int", np.full((5, 5), 0.5))[0] == _lib.SPM_ENOTINT
        # |T| bound: 2^27 * prod(row sums 28 * 2^20) >= 2^126
        assert eng.try_submit("int", np.full((28, 28), float(1 << 20)))[0] == _lib.SPM_ERANGE


def test_request_budget_splits_batches():
    with BatchEngine(dry_run=True, max_requests=3) as eng:
        st, res = eng.run_raw(reqs(10))
        assert st == 0
    with BatchEngine(dry_run=True, max_requests=3) as eng:
        rc, res, lat, stats = eng.replay(reqs(10), np.zeros(10))
        assert rc == 0 and stats.batches >= 4 and stats.mean_batch <= 3


def test_work_budget_one_request_per_batch():
    with BatchEngine(dry_run=True, max_batch_work=1.0) as eng:
        rc, res, lat, stats = eng.replay(reqs(12), np.zeros(12))
        assert rc == 0 and stats.batches == 12


def test_item_budget():
    # n = 20 real with 2^10-step chunks: 2^19 / 2^10 / 64 = 8 items each
    with BatchEngine(dry_run=True, max_items=16) as eng:
        r = [("real", np.ones((20, 20)), i) for i in range(6)]
        rc, res, lat, stats = eng.replay(r, np.zeros(6))
        assert rc == 0 and stats.batches >= 3


def test_cancel_queued_and_unknown():
    with BatchEngine(dry_run=True, slots=1, max_requests=1) as eng:
        big = np.ones((28, 28))
        t1 = eng.submit("complex", big)   # ~1.5 ms of simulated work
        t2 = eng.submit("real", np.ones((8, 8)))
        t3 = eng.submit("real", np.ones((8, 8)))
        eng.progress()                    # t1 launched, t2/t3 queued
        assert eng.cancel(t2) == 0
        assert eng.cancel(123456) == _lib.SPM_EINVAL
        got = {c[0]: c[2] for c in drain(eng)}
        assert got[t1] == 0 and got[t2] == _lib.SPM_ECANCELLED and got[t3] == 0


def test_cancel_running():
    with BatchEngine(dry_run=True, slots=1) as eng:
        t1 = eng.submit("complex", np.ones((28, 28)))
        eng.progress()
        assert eng.pending() == (0, 1)
        assert eng.cancel(t1) == 0
        got = drain(eng)
        assert got[0][0] == t1 and got[0][2] == _lib.SPM_ECANCELLED


def test_deadline_expires_while_queued():
    with BatchEngine(dry_run=True, slots=1, max_requests=1) as eng:
        t1 = eng.submit("complex", np.ones((28, 28)))
        eng.progress()
        t2 = eng.submit("real", np.ones((8, 8)), deadline_s=1e-4)
        time.sleep(3e-3)
        got = {c[0]: c[2] for c in drain(eng)}
        assert got[t1] == 0 and got[t2] == _lib.SPM_ETIMEDOUT


def test_injected_fault_fails_one_batch_only():
    with BatchEngine(dry_run=True, max_requests=4) as eng:
        eng.inject_fault(1)
        st, res = eng.run_raw(reqs(8))
        assert st == _lib.SPM_EFAULTINJ
        statuses = [r.status for r in res]
        assert statuses.count(_lib.SPM_EFAULTINJ) == 4 and statuses.count(0) == 4
        st, res = eng.run_raw(reqs(8))  # the engine keeps serving
        assert st == 0


def test_results_failing_their_checks_run_again():
    with BatchEngine(dry_run=True, max_requests=4) as eng:
        eng.inject_bad_results(3)
        rc, res, _, st = eng.replay(reqs(8), np.zeros(8))
        assert rc == 0 and all(r.status == 0 for r in res)
        assert st.retried == 3 and st.completed == 8
        eng.inject_bad_results(100)  # a request that never passes: two re-runs, then an error
        rc, res, _, st = eng.replay(reqs(1), np.zeros(1))
        assert rc == _lib.SPM_EHIP and res[0].status == _lib.SPM_EHIP and st.retried == 2
        assert eng.last_error() == "injected bad result"
        eng.inject_bad_results(0)
        rc, res, _, st = eng.replay(reqs(4), np.zeros(4))
        assert rc == 0 and st.retried == 0


def test_replay_latency_accounting():
    stream = W.serving_stream(200, rate=2e4, seed=5, n_lo=8, n_hi=16)
    r = [(q.kind, q.a, q.rid) for q in stream]
    arr = np.array([q.arrival for q in stream])
    with BatchEngine(dry_run=True) as eng:
        rc, res, lat, stats = eng.replay(r, arr)
    assert rc == 0 and stats.completed == 200 and stats.failed == 0
    assert (lat >= 0).all() and stats.p50_s <= stats.p99_s <= stats.max_s
    assert stats.wall_s >= arr[-1] - arr[0]


def test_replay_rejects_unsorted_arrivals():
    with BatchEngine(dry_run=True) as eng:
        with pytest.raises(Exception):
            rc, *_ = eng.replay(reqs(3), np.array([0.0, 0.2, 0.1]))
            _lib.check(rc)


def test_tuning_roundtrip_and_validation():
    with BatchEngine(dry_run=True) as eng:
        t = eng.tuning
        t.fused, t.waves_per_cu = True, 8
        t.log2_chunk["complex"] = 7
        eng.tuning = t
        assert eng.tuning == t
        t.log2_chunk["real"] = 1
        with pytest.raises(_lib.SpmError):
            eng.tuning = t
