"""Benchmark: one exact permanent of a dense real FP64 40 x 40 matrix with
i.i.d. N(0,1) entries (seed 40), compensated (TwoSum) accumulation -- 2^39
Gray-code subsets per step -- on one MI355X, through the C ABI with the matrix
resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision kahan|dd]
                    [--no-cpu-baseline] [--n 40]

Prints ONE JSON line.  value = permanents per second (1 / seconds per
permanent).  roofline.achieved counts the algorithmic 2n FP64 flops per
subset (n adds + n-1 multiplies + 1 accumulate) over the kernel's duration
measured with HIP events on its own stream; peak = MI355X FP64 vector.
cpu_baseline = the reference's CPU Gray-code walk (restated in oracle/, same
chunking and arithmetic order) on a bounded sample of the same walk, timed
on this host's cores and scaled to permanents per second.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "dense n=40 perm time & % FP64 VALU peak; batched perms/s + p99 latency"
PEAK_FP64_TFLOPS = 78.6


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(a: np.ndarray, target_s: float = 12.0) -> dict:
    from oracle import oracle as O  # the checker / CPU baseline only
    n = a.shape[0]
    cores = host_cores()
    count = 1 << 22
    t0 = time.perf_counter()
    O.dense_ref_sample(a, cores, count)
    dt = time.perf_counter() - t0
    count = int(min(1 << (n - 1), max(count, count * target_s / max(dt, 1e-6))))
    t0 = time.perf_counter()
    O.dense_ref_sample(a, cores, count)
    dt = time.perf_counter() - t0
    rate = count / dt  # subsets per second
    return {"value": rate / float(1 << (n - 1)), "unit": "perm/s", "cores": cores, "kind": "port",
            "sample": f"reference dpNaivePerman walk (oracle/ryser_oracle.c orc_dense_ref_sample) over "
                      f"Gray indices [1, 1+{count}) of the same n={n} matrix on {cores} threads, {dt:.1f} s; "
                      f"{rate:.3e} subsets/s scaled to 2^{n - 1} subsets per permanent"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="kahan", choices=["kahan", "dd"])
    ap.add_argument("--n", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the permanent's 2^(n-1) subset range would shard "
              "across GPUs with one 16-byte partial sum each, but multi-GPU runs are not part of "
              "this build (use --gpus 1).", file=sys.stderr)
        return 2

    import superman_amd as S
    from superman_amd import workloads as W

    n = args.n
    a = W.dense_normal(n, 40 if n == 40 else n)
    dev = S.device_info(0)
    plan = S.DensePlan(a, args.precision)
    info = plan.info()

    for _ in range(args.warmup):
        plan.run()
    S.synchronize(0)
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = plan.run()
        kernel_ms.append(r.kernel_ms)
    S.synchronize(0)
    dt = time.perf_counter() - t0
    plan.close()
    # the other accumulation mode, once, outside the timed region
    other = "dd" if args.precision == "kahan" else "kahan"
    with S.DensePlan(a, other) as p2:
        r_other = p2.run()

    ms_per_step = dt * 1e3 / args.steps
    flops = 2.0 * n * float(1 << (n - 1))  # algorithmic 2n flops per subset
    k_ms = float(np.mean(kernel_ms))
    achieved = flops / (k_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": 1e3 / ms_per_step,
        "unit": f"perm/s (dense FP64 n={n})",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"dense real FP64 n={n}, entries i.i.d. N(0,1) from numpy default_rng(40); "
                               f"one permanent per step ({'compensated TwoSum' if args.precision == 'kahan' else 'plain'} "
                               f"accumulation), 2^{n - 1} Gray-code subsets",
                   "n": n, "precision": args.precision, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": info["kernel"], "kernel_ms": k_ms,
                     "flops_per_launch": flops, "flop_convention": "2n per subset (add, mul, accumulate; FMA = 2)"},
        "result": {"value": r.value, "err_est": r.err_est, "cond": r.cond,
                   f"{other}_value": r_other.value, f"{other}_kernel_ms": r_other.kernel_ms},
        "geometry": info,
        "device": dev,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
