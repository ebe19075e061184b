"""Synthetic inputs for the permanent workloads (seeded numpy PCG64).

These are the matrices the benchmark and the parity tests run on:

  dense_uniform(n, seed)   entries i.i.d. uniform [0, 1)
  dense_normal(n, seed)    entries i.i.d. N(0, 1) (signed: heavy cancellation)
  sparse01(n, seed)        0/1, row i holds p[i] of a random permutation p plus
                           3 further distinct random columns (~4 nnz per row,
                           a perfect matching is guaranteed)
  haar_block(n, seed, m)   top-left n x n block of a Haar-random m x m unitary
                           (QR of an i.i.d. complex Gaussian matrix with the
                           diagonal phases of R folded into Q)
  complex_gauss(n, seed)   entries i.i.d. standard complex Gaussian
  serving_stream(...)      Poisson request stream with mixed kinds and sizes
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def dense_uniform(n: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).random((n, n))


def dense_normal(n: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal((n, n))


def sparse01(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = rng.permutation(n)
    a = np.zeros((n, n))
    for i in range(n):
        a[i, p[i]] = 1.0
        others = [c for c in range(n) if c != p[i]]
        a[i, rng.choice(others, 3, replace=False)] = 1.0
    return a


def haar_block(n: int, seed: int = 32, m: int = 256) -> np.ndarray:
    rng = np.random.default_rng(seed)
    z = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    q = q * (d / np.abs(d))[None, :]
    return np.ascontiguousarray(q[:n, :n])


def complex_gauss(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2.0)


@dataclass
class Request:
    rid: int
    arrival: float          # seconds from stream start
    kind: str               # "real" | "complex" | "int01"
    a: np.ndarray


def serving_stream(count: int, rate: float, seed: int = 2026, n_lo: int = 8, n_hi: int = 28):
    """Requests with exponential inter-arrival times (rate per second), n
    uniform in [n_lo, n_hi]; 50% real uniform [0,1), 30% complex Gaussian,
    20% dense 0/1 with density 0.5."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(count):
        t += rng.exponential(1.0 / rate)
        n = int(rng.integers(n_lo, n_hi + 1))
        u = rng.random()
        if u < 0.5:
            kind, a = "real", rng.random((n, n))
        elif u < 0.8:
            kind = "complex"
            a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2.0)
        else:
            kind, a = "int01", (rng.random((n, n)) < 0.5).astype(np.float64)
        out.append(Request(rid, t, kind, a))
    return out
