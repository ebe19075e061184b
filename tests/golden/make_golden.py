"""Writes the golden fixtures in tests/golden/ (run in the development container).

    python tests/golden/make_golden.py [--skip-slow]

Fixtures are data only: each input matrix is saved as .npy next to
golden.json, which records for it
  * ref1 / ref8   -- the output of the reference's own CPU path
                     (oracle/_ref/superman_ref_cpu, compiled from the
                     reference sources by `make -C oracle ref`) at
                     thread_count=1 (deterministic) and 8, with its wall time;
  * dd            -- the double-double oracle value (hi, lo);
  * exact         -- the exact-integer oracle value (decimal string);
  * known         -- a closed-form answer, where one exists.
The reference binary only exists where the reference sources do; the GPU
tests read golden.json and the .npy files only.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oracle import refbin as R  # noqa: E402
from superman_amd import workloads as W  # noqa: E402


def save(name: str, a: np.ndarray) -> str:
    fn = name + ".npy"
    np.save(os.path.join(HERE, fn), a, allow_pickle=False)
    return fn


def cval(v):
    if isinstance(v, complex):
        return [v.real, v.imag]
    return v


def ref_runs(a, threads=(1, 8), reps8=1, timeout=3600):
    out = {}
    for t in threads:
        vals, times = [], []
        for _ in range(reps8 if t > 1 else 1):
            v, s, _ = R.run(a, t, timeout=timeout)
            vals.append(v)
            times.append(s)
        out[f"ref{t}"] = cval(vals[0])
        out[f"ref{t}_seconds"] = statistics.median(times)
        if len(vals) > 1:
            out[f"ref{t}_all"] = [cval(v) for v in vals]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-slow", action="store_true")
    args = ap.parse_args()
    if not R.available():
        raise SystemExit("build the reference CPU binary first: make -C oracle ref")
    O.build()
    cases = []

    def add(name, kind, a, **vals):
        e = {"name": name, "kind": kind, "n": int(a.shape[0]), "file": save(name, a)}
        e.update(vals)
        cases.append(e)
        print(name, {k: v for k, v in e.items() if k not in ("name", "file")}, flush=True)

    # ---- known answers ----------------------------------------------------
    a = np.array([[1, 2, 3], [4, 5, 6], [7, 8, 10]], dtype=np.float64)
    add("kat_3x3_463", "int", a, known="463", exact=str(O.int_exact(a.astype(np.int64))[0]), **ref_runs(a))
    a = np.ones((12, 12))
    add("kat_ones12", "int", a, known=str(math.factorial(12)), **ref_runs(a))
    a = np.full((10, 10), 1j)
    add("kat_all_i10", "complex", a, known=[-float(math.factorial(10)), 0.0], **ref_runs(a))

    # ---- dense real: reference anchors + dd oracle ------------------------
    for n in (1, 2, 3, 5, 8, 13, 16, 20, 24):
        a = W.dense_uniform(n, n)
        hi, lo = O.dense_dd(a)
        add(f"dense_uniform_n{n}_s{n}", "real", a, dd=[hi, lo], **ref_runs(a))
    for n in (4, 9, 16, 20, 24, 28):
        a = W.dense_normal(n, n)
        hi, lo = O.dense_dd(a)
        add(f"dense_normal_n{n}_s{n}", "real", a, dd=[hi, lo], **ref_runs(a))
    # config 1: dense uniform n=24, seed 24 (reference timing: median of 3)
    a = W.dense_uniform(24, 24)
    hi, lo = O.dense_dd(a)
    add("c1_dense_uniform_n24_s24", "real", a, dd=[hi, lo], **ref_runs(a, reps8=3))
    # config 2 oracle sizes (seed = n) and the n=40 inputs
    for n in (28, 30):
        a = W.dense_uniform(n, n)
        hi, lo = O.dense_dd(a)
        add(f"dense_uniform_n{n}_s{n}", "real", a, dd=[hi, lo])
    a = W.dense_normal(30, 30)
    hi, lo = O.dense_dd(a)
    add("dense_normal_n30_s30", "real", a, dd=[hi, lo])
    add("c2_dense_normal_n40_s40", "real", W.dense_normal(40, 40))
    add("c2_dense_uniform_n40_s40", "real", W.dense_uniform(40, 40))

    # ---- complex ------------------------------------------------------------
    for n in (3, 7, 12, 16, 20):
        a = W.complex_gauss(n, n)
        d = O.complex_dd(a)
        add(f"complex_gauss_n{n}_s{n}", "complex", a, dd=list(d), **ref_runs(a))
    for n in (16, 20, 24):
        a = W.haar_block(n, 32)
        d = O.complex_dd(a)
        add(f"haar_n{n}_s32", "complex", a, dd=list(d), **ref_runs(a))
    a = W.haar_block(28, 32)
    add("haar_n28_s32", "complex", a, dd=list(O.complex_dd(a)))
    a = W.haar_block(32, 32)
    if args.skip_slow:
        add("c4_haar_n32_s32", "complex", a)
    else:
        t = time.time()
        d = O.complex_dd(a)
        add("c4_haar_n32_s32", "complex", a, dd=list(d), dd_seconds=time.time() - t)

    # ---- integer / sparse ---------------------------------------------------
    for n, s in ((12, 12), (16, 16), (20, 20), (24, 24), (30, 7)):
        a = W.sparse01(n, s)
        v, st = O.int_exact(a.astype(np.int64))
        add(f"sparse01_n{n}_s{s}", "int", a, exact=str(v), exact_status=st, **ref_runs(a))
    for s in range(40, 45):
        a = W.sparse01(40, s)
        r = ref_runs(a, threads=(8,))
        add(f"c3_sparse01_n40_s{s}", "int", a, **r)
    for n in (8, 12, 16, 20, 24):
        a = (np.random.default_rng(n).random((n, n)) < 0.5).astype(np.float64)
        v, st = O.int_exact(a.astype(np.int64))
        add(f"int01_d50_n{n}_s{n}", "int", a, exact=str(v), exact_status=st, **ref_runs(a))
    a = (np.random.default_rng(28).random((28, 28)) < 0.5).astype(np.float64)
    v, st = O.int_exact(a.astype(np.int64))
    add("int01_d50_n28_s28", "int", a, exact=str(v), exact_status=st)

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "numpy": np.__version__,
                   "cases": cases}, f, indent=1)
    print("wrote", len(cases), "cases")


if __name__ == "__main__":
    main()
