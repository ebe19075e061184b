"""Batched and served permanents (spm_batch_* / spm_serve_* of the C ABI).

    with BatchEngine() as eng:
        results = eng.run([("real", a1), ("complex", z), ("int", b)])

A request is (kind, matrix) with kind "real" (compensated FP64), "complex"
(compensated complex128) or "int" (exact), 1 <= n <= 28.  The engine keeps
`slots` batches in flight; each batch is one persistent launch over a device
work queue.  Results come back as api.Permanent.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import SpmBatchConfig, SpmBatchTuning, SpmCompletion, SpmRequest, SpmResult, SpmServeStats, check, lib
from .api import Permanent, _to_perm

KINDS = {"real": _lib.SPM_KIND_REAL, "complex": _lib.SPM_KIND_COMPLEX, "int": _lib.SPM_KIND_INT,
         "int01": _lib.SPM_KIND_INT}
KIND_NAMES = {_lib.SPM_KIND_REAL: "real", _lib.SPM_KIND_COMPLEX: "complex", _lib.SPM_KIND_INT: "int"}
MAX_N = _lib.SPM_BATCH_MAX_N


def kind_of(k) -> int:
    if isinstance(k, (int, np.integer)):
        return int(k)
    try:
        return KINDS[k]
    except KeyError:
        raise ValueError(f"unknown request kind {k!r}") from None


@dataclass
class Tuning:
    fused: bool
    waves_per_cu: int
    log2_chunk: dict

    def to_c(self) -> SpmBatchTuning:
        t = SpmBatchTuning()
        t.fused = int(bool(self.fused))
        t.waves_per_cu = int(self.waves_per_cu)
        for k, v in self.log2_chunk.items():
            t.log2_chunk[kind_of(k)] = int(v)
        return t

    @staticmethod
    def from_c(t: SpmBatchTuning) -> "Tuning":
        return Tuning(bool(t.fused), int(t.waves_per_cu), {KIND_NAMES[k]: int(t.log2_chunk[k]) for k in (1, 2, 3)})


@dataclass
class ServeStats:
    completed: int
    failed: int
    batches: int
    wall_s: float
    perms_per_s: float
    p50_s: float
    p99_s: float
    mean_s: float
    max_s: float
    mean_batch: float
    device_busy: float
    late_publish: int = 0
    retried: int = 0


class _Requests:
    """ctypes request array plus the matrix buffers it points into."""

    def __init__(self, reqs, deadline_s: float = 0.0):
        self.bufs = []
        self.kinds = []
        self.arr = (SpmRequest * max(1, len(reqs)))()
        for i, req in enumerate(reqs):
            kind, a = req[0], req[1]
            k = kind_of(kind)
            if k == _lib.SPM_KIND_COMPLEX:
                buf = np.ascontiguousarray(a, dtype=np.complex128)
                dptr = buf.view(np.float64)
            else:
                buf = np.ascontiguousarray(a, dtype=np.float64)
                dptr = buf
            if buf.ndim != 2 or buf.shape[0] != buf.shape[1]:
                raise ValueError("square matrix expected")
            self.bufs.append(dptr)
            self.kinds.append(k)
            self.arr[i].kind = k
            self.arr[i].n = buf.shape[0]
            self.arr[i].a = dptr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            self.arr[i].tag = req[2] if len(req) > 2 else i
            self.arr[i].deadline_s = deadline_s
        self.count = len(reqs)


def _perm(r: SpmResult, kind: int) -> Permanent:
    return _to_perm(r, kind == _lib.SPM_KIND_COMPLEX)


class BatchEngine:
    def __init__(self, device: int = 0, slots: int = 2, max_requests: int = 1024, max_items: int = 1 << 16,
                 arena_bytes: int = 64 << 20, max_pending: int = 1 << 16, max_batch_work: float = 0.0,
                 use_graph: bool = True, dry_run: bool = False):
        cfg = SpmBatchConfig()
        lib().spm_batch_config_default(ctypes.byref(cfg))
        cfg.device, cfg.slots, cfg.max_requests, cfg.max_items = device, slots, max_requests, max_items
        cfg.arena_bytes, cfg.max_pending, cfg.max_batch_work = arena_bytes, max_pending, max_batch_work
        cfg.use_graph = int(use_graph)
        cfg.executor = _lib.SPM_EXEC_DRYRUN if dry_run else _lib.SPM_EXEC_DEVICE
        self._h = ctypes.c_void_p()
        check(lib().spm_batch_create(ctypes.byref(cfg), ctypes.byref(self._h)), "spm_batch_create")
        self._live = {}  # ticket -> (kind, buffers) for served requests

    def close(self):
        if self._h:
            lib().spm_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- tuning ------------------------------------------------------------
    @property
    def tuning(self) -> Tuning:
        t = SpmBatchTuning()
        check(lib().spm_batch_get_tuning(self._h, ctypes.byref(t)))
        return Tuning.from_c(t)

    @tuning.setter
    def tuning(self, t: Tuning) -> None:
        c = t.to_c()
        check(lib().spm_batch_set_tuning(self._h, ctypes.byref(c)), "spm_batch_set_tuning")

    def autotune(self, sample, reps: int = 2) -> tuple[Tuning, float]:
        rq = _Requests(sample)
        best = SpmBatchTuning()
        sec = ctypes.c_double()
        st = lib().spm_batch_autotune(self._h, rq.count, rq.arr, reps, ctypes.byref(best), ctypes.byref(sec))
        check(st, "spm_batch_autotune" + (f" ({self.last_error()})" if st == _lib.SPM_EHIP else ""))
        return Tuning.from_c(best), sec.value

    # -- synchronous batches -----------------------------------------------
    def run_raw(self, reqs) -> tuple[int, list]:
        """(first error status, [SpmResult]) without raising."""
        rq = _Requests(reqs)
        out = (SpmResult * max(1, rq.count))()
        st = lib().spm_batch_run(self._h, rq.count, rq.arr, out)
        return st, [out[i] for i in range(rq.count)]

    def run(self, reqs) -> list[Permanent]:
        st, res = self.run_raw(reqs)
        check(st, "spm_batch_run" + (f" ({self.last_error()})" if st == _lib.SPM_EHIP else ""))
        return [_perm(r, kind_of(q[0])) for r, q in zip(res, reqs)]

    # -- serving ------------------------------------------------------------
    def submit(self, kind, a, tag: int = 0, deadline_s: float = 0.0) -> int:
        rq = _Requests([(kind, a, tag)], deadline_s)
        tk = ctypes.c_uint64()
        check(lib().spm_serve_submit(self._h, rq.arr, ctypes.byref(tk)), "spm_serve_submit")
        self._live[tk.value] = rq.kinds[0]
        return tk.value

    def try_submit(self, kind, a, tag: int = 0, deadline_s: float = 0.0) -> tuple[int, int]:
        """(status, ticket): admission refusals come back as a status."""
        rq = _Requests([(kind, a, tag)], deadline_s)
        tk = ctypes.c_uint64()
        st = lib().spm_serve_submit(self._h, rq.arr, ctypes.byref(tk))
        if st == _lib.SPM_OK:
            self._live[tk.value] = rq.kinds[0]
        return st, tk.value

    def progress(self) -> int:
        k = ctypes.c_int()
        check(lib().spm_serve_progress(self._h, ctypes.byref(k)))
        return k.value

    def poll(self, max_items: int = 256) -> list[tuple[int, int, int, Permanent | None, SpmCompletion]]:
        """[(ticket, tag, status, Permanent or None, raw completion)]."""
        buf = (SpmCompletion * max_items)()
        got = ctypes.c_int()
        check(lib().spm_serve_poll(self._h, buf, max_items, ctypes.byref(got)))
        out = []
        for i in range(got.value):
            c = buf[i]
            kind = self._live.pop(c.ticket, _lib.SPM_KIND_REAL)
            p = _perm(c.result, kind) if c.result.status == _lib.SPM_OK else None
            out.append((c.ticket, c.tag, c.result.status, p, c))
        return out

    def cancel(self, ticket: int) -> int:
        return lib().spm_serve_cancel(self._h, ticket)

    def pending(self) -> tuple[int, int]:
        q, f = ctypes.c_int(), ctypes.c_int()
        check(lib().spm_serve_pending(self._h, ctypes.byref(q), ctypes.byref(f)))
        return q.value, f.value

    def last_error(self) -> str:
        return lib().spm_batch_last_error(self._h).decode()

    def counters(self) -> dict:
        """Engine-lifetime late_publish / retried totals (0 on a healthy run)."""
        late, retried = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().spm_batch_counters(self._h, ctypes.byref(late), ctypes.byref(retried)))
        return {"late_publish": late.value, "retried": retried.value}

    def inject_fault(self, launches: int = 1) -> None:
        check(lib().spm_batch_inject_fault(self._h, launches))

    def inject_bad_results(self, results: int = 1) -> None:
        check(lib().spm_batch_inject_bad_results(self._h, results))

    def replay(self, reqs, arrival_s) -> tuple[int, list, np.ndarray, ServeStats]:
        """Replays an open-loop arrival schedule; (status, [SpmResult], latency_s, stats)."""
        rq = _Requests(reqs)
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        if arr.shape != (rq.count,):
            raise ValueError("one arrival time per request")
        out = (SpmResult * max(1, rq.count))()
        lat = np.zeros(rq.count)
        st = SpmServeStats()
        rc = lib().spm_serve_replay(self._h, rq.count, rq.arr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    out, lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(st))
        stats = ServeStats(int(st.completed), int(st.failed), int(st.batches), st.wall_s, st.perms_per_s, st.p50_s,
                           st.p99_s, st.mean_s, st.max_s, st.mean_batch, st.device_busy, int(st.late_publish),
                           int(st.retried))
        return rc, [out[i] for i in range(rq.count)], lat, stats


def to_permanent(r: SpmResult, kind) -> Permanent:
    return _perm(r, kind_of(kind))
