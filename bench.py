"""Benchmark: one exact permanent of a dense real FP64 40 x 40 matrix with
i.i.d. N(0,1) entries (seed 40), compensated (TwoSum) accumulation -- 2^39
Gray-code subsets per step -- on one MI355X, through the C ABI with the matrix
resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision kahan|dd]
                    [--no-cpu-baseline] [--n 40] [--serving-requests 20000]
                    [--no-serving] [--no-secondary]

Prints ONE JSON line.  value = permanents per second (1 / seconds per
permanent).  The line also carries the other halves of the metric:
`serving` -- the 20,000-request Poisson stream (n uniform in 8..28; 50% real
uniform [0,1), 30% complex Gaussian, 20% 0/1 density 0.5, seed 2026) through
the batched persistent kernel (autotuned, hipGraph launches, continuous
batching) at 80% of its measured capacity, with p50/p99 latency, against the
plain one-call-per-matrix GPU path on the same requests, with the speed-up
that removing every per-call cost would give (`compute_bound`) and both
paths' latency at one offered load the plain path sustains (`equal_load`);
`serving_small_n` -- the same for orders 8..20, where a plain call is
mostly per-call cost; and `secondary` -- the dense n=24, sparse 0/1 n=40
(generated kernel) and Haar complex n=32 configurations, one permanent each
with inputs resident.

roofline.achieved counts the algorithmic 2n FP64 flops per subset (n adds +
n-1 multiplies + 1 accumulate) over the kernel's duration measured with HIP
events on its own stream; peak = MI355X FP64 vector; ceiling = the issue
ceiling 2n/(2(2n-1)) of peak of any Gray-code Ryser (2n-1 FP64 instructions
per subset).
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

# The generated kernel's compile time is reported as a compile: this process
# gets a fresh, private kernel cache and no ROCm code-object cache (set
# before the library or hiprtc is loaded).
if "SPM_KERNEL_CACHE" not in os.environ:
    import atexit
    import shutil
    import tempfile
    os.environ["SPM_KERNEL_CACHE"] = tempfile.mkdtemp(prefix="spm_bench_kernels_")
    atexit.register(shutil.rmtree, os.environ["SPM_KERNEL_CACHE"], True)
os.environ.setdefault("AMD_COMGR_CACHE", "0")

import numpy as np  # noqa: E402

METRIC = "dense n=40 perm time & % FP64 VALU peak; batched perms/s + p99 latency"
PEAK_FP64_TFLOPS = 78.6


def host_threads() -> int:
    """Threads for the CPU baseline: OMP_NUM_THREADS where the host sets it
    (the CPU share that goes with one GPU), else every CPU this process may
    run on."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return max(1, min(int(env), aff))
    return max(1, aff)


def host_info() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = None
    return {"nproc": os.cpu_count(), "affinity_cpus": aff, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "threads_used": host_threads(), "cpu_model": model,
            "thread_rule": "OMP_NUM_THREADS when set (the CPU share of one GPU on a shared host), else all CPUs "
                           "in this process's affinity mask"}


def _median_time(fn, runs: int = 3):
    ts, v = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        v = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), v


def _sized_sample(fn, n: int, target_s: float) -> tuple[int, float]:
    """Runs fn(count) on a short prefix, then on a prefix sized to about
    target_s seconds (at most the whole 2^(n-1) range); (count, seconds)."""
    count = 1 << 20
    t0 = time.perf_counter()
    fn(count)
    dt = time.perf_counter() - t0
    count = int(min((1 << (n - 1)) - 1, max(count, count * target_s / max(dt, 1e-6))))
    t0 = time.perf_counter()
    fn(count)
    return count, time.perf_counter() - t0


def cpu_baseline(a: np.ndarray, target_s: float = 12.0) -> dict:
    """The reference's CPU algorithms on this host: the headline n=40 dense
    permanent as a bounded sample of the restated walk (a whole one takes
    hours), the n=24 dense (whole), complex Haar n=32 (sample) and sparse
    0/1 n=40 (whole, decomposed by the reference itself) configurations."""
    from oracle import oracle as O  # CPU baseline timing only
    from oracle import refbin
    from superman_amd import workloads as W
    n = a.shape[0]
    threads = host_threads()
    count, dt = _sized_sample(lambda c: O.dense_ref_sample(a, threads, c), n, target_s)
    rate = count / dt  # subsets per second
    out = {"value": rate / float(1 << (n - 1)), "unit": "perm/s", "cores": threads, "kind": "port",
           "sample": f"reference dpNaivePerman walk (oracle/ryser_oracle.c orc_dense_ref_sample) over "
                     f"Gray indices [1, 1+{count}) of the same n={n} matrix on {threads} threads, {dt:.1f} s; "
                     f"{rate:.3e} subsets/s ({rate / threads:.3e} per thread) scaled to 2^{n - 1} subsets "
                     f"per permanent",
           "host": host_info()}
    have_ref = refbin.available()
    # config 1: dense uniform n=24, one whole permanent, median of 3
    c1 = W.dense_uniform(24, 24)
    port_s, port_v = _median_time(lambda: O.dense_ref(c1, threads))
    d1 = {"port_s_median": port_s, "port_value": port_v, "threads": threads, "runs": 3}
    if have_ref:
        rts = [refbin.run(c1, threads=threads) for _ in range(3)]
        d1.update({"reference_s_median": float(np.median([r[1] for r in rts])), "reference_value": rts[-1][0],
                   "reference_timer": "the binary's own 'Computed in' seconds (main.cpp:122-131)"})
    out["c1_dense_uniform_n24"] = d1
    # config 4: Haar complex n=32, bounded sample of the restated complex walk
    z = W.haar_block(32, 32)
    zc, zdt = _sized_sample(lambda c: O.complex_ref_sample(z, threads, c), 32, 5.0)
    out["c4_complex_haar_n32"] = {
        "perm_per_s": zc / zdt / float(1 << 31), "seconds_per_perm": float(1 << 31) * zdt / zc,
        "kind": "port", "threads": threads,
        "sample": f"cpuComputeComplex walk (orc_complex_ref_sample) over Gray indices [1, 1+{zc}) on "
                  f"{threads} threads, {zdt:.1f} s, scaled to 2^31 subsets"}
    # config 3: sparse 0/1 n=40 through the reference's own pipeline (DM
    # pruning, components, compressions, CPU sparse walk), median of 3
    if have_ref:
        b = W.sparse01(40, 40)
        rts = [refbin.run(b, threads=threads) for _ in range(3)]
        out["c3_sparse01_n40"] = {"reference_s_median": float(np.median([r[1] for r in rts])),
                                  "reference_value": rts[-1][0], "kind": "reference", "threads": threads}
    rc = reference_check(threads)
    out["reference_check"] = rc
    if rc is not None:
        # the reference binary's complex walk is slower than the port (its
        # std::complex products go through the C runtime's NaN-checking
        # multiply), so config 4's baseline is the binary's whole-permanent
        # rate at n=24 scaled by subsets x order to n=32
        c = rc["complex"]
        t32 = c["reference_s"] * float(1 << (31 - (c["n"] - 1))) * 32.0 / c["n"]
        out["c4_complex_haar_n32"].update({"reference_seconds_per_perm_extrapolated": t32,
                                           "reference_perm_per_s_extrapolated": 1.0 / t32})
    return out


def reference_check(cores: int, n: int = 30, nc: int = 24) -> dict | None:
    """The reference's own CPU binary (oracle/_ref, built from its sources)
    and the port on one whole n=30 dense permanent and one whole n=24 complex
    permanent of the same generators, same threads: the port's speed stands
    in for the reference's on the n=40 and complex n=32 samples."""
    from oracle import oracle as O
    from oracle import refbin
    if not refbin.available():
        return None
    from superman_amd import workloads as W
    a = W.dense_normal(n, n)
    t0 = time.perf_counter()
    v_port = O.dense_ref(a, cores)
    t_port = time.perf_counter() - t0
    v_ref, t_ref, _ = refbin.run(a, threads=cores)
    sub = float(1 << (n - 1))
    z = W.complex_gauss(nc, nc)
    t0 = time.perf_counter()
    z_port = O.complex_ref(z, cores)
    tz_port = time.perf_counter() - t0
    z_ref, tz_ref, _ = refbin.run(z, threads=cores)
    subc = float(1 << (nc - 1))
    return {"n": n, "threads": cores, "reference_s": t_ref, "port_s": t_port,
            "reference_subsets_per_s": sub / t_ref, "port_subsets_per_s": sub / t_port,
            "values_rel_diff": abs(v_port - v_ref) / abs(v_ref),
            "complex": {"n": nc, "reference_s": tz_ref, "port_s": tz_port,
                        "reference_subsets_per_s": subc / tz_ref, "port_subsets_per_s": subc / tz_port,
                        "values_rel_diff": abs(z_port - z_ref) / abs(z_ref)}}


def measured_traffic(kernel: str):
    """HBM bytes per launch from the committed counter runs (two rocprofv3
    --pmc passes, FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of known
    size), when they were taken on this kernel; else None."""
    path = os.path.join(ROOT, "profiles", "dense_f64_n40_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except OSError:
        return None
    return t["bytes_per_launch"] if t.get("kernel") == kernel else None


def measured_counters(kernel: str, kernel_ms: float):
    """Counter-based figures for the dominant kernel from the committed PMC
    run (instruction and FP64-flop counts per launch do not depend on the
    box), turned into rates with this run's kernel time."""
    path = os.path.join(ROOT, "profiles", "dense_f64_n40_counters.json")
    try:
        with open(path) as f:
            c = json.load(f)
    except OSError:
        return None
    if c.get("kernel") != kernel or kernel_ms <= 0:
        return None
    t = kernel_ms * 1e-3
    counted = c["counted_fp64_flops_per_launch"] / t / 1e12
    issue = c["sq_insts_valu_per_launch"] * 4.0 / (1024 * 2.4e9 * t)  # 4 cycles per wave64 FP64 op, 1024 SIMDs
    return {"counted_tflops": counted, "counted_frac": counted / PEAK_FP64_TFLOPS,
            "counted_rule": c["counted_flops_rule"], "valu_issue_frac": issue,
            "valu_busy_pmc": c["valu_busy"], "source": c["source"]}


def request_flops(kind: str, n: int) -> float:
    """Algorithmic FP64 work of one request: 2n flops per Gray subset (real
    and the exact FP64 walk of 0/1 matrices), 8n (complex)."""
    return (8.0 if kind == "complex" else 2.0) * n * float(1 << (n - 1))


def plain_replay(reqs, arrivals, calls) -> np.ndarray:
    """One C-ABI call per request, each started at its arrival time (or as
    soon as the previous call returns): arrival -> result latencies."""
    lat = np.zeros(len(reqs))
    t0 = time.perf_counter()
    for i, ((kind, a, _), t_arr) in enumerate(zip(reqs, arrivals)):
        while time.perf_counter() - t0 < t_arr:
            pass
        calls[kind](a)
        lat[i] = time.perf_counter() - t0 - t_arr
    return lat


def serving(count: int, seed: int = 2026, load: float = 0.8, n_lo: int = 8, n_hi: int = 28,
            valu_tflops: float | None = None, equal_load_count: int = 5000) -> dict:
    """The serving stream through the batch engine (capacity, then Poisson
    arrivals at `load` of it) against one C-ABI call per matrix.
    `speedup_bound` = the plain path's wall time over the stream's
    algorithmic work at `valu_tflops` (the dense n=40 kernel's measured
    rate): what batching could reach if it removed every per-call cost and
    ran every request at that rate."""
    from superman_amd import api
    from superman_amd import workloads as W
    from superman_amd.batch import BatchEngine, to_permanent

    stream = W.serving_stream(count, rate=1.0, seed=seed, n_lo=n_lo, n_hi=n_hi)
    reqs = [(q.kind, q.a, q.rid) for q in stream]
    unit_arrivals = np.array([q.arrival for q in stream])  # rate 1/s; scaled to the offered rate below
    with BatchEngine() as eng:
        tuning, _ = eng.autotune(reqs[:1000], reps=1)
        rc, res_sat, _, sat = eng.replay(reqs, np.zeros(count))  # capacity: everything queued at t = 0
        if rc != 0:
            raise RuntimeError(f"batched capacity run failed: status {rc} ({eng.last_error()})")
        rate = load * sat.perms_per_s
        rc, res, lat, st = eng.replay(reqs, unit_arrivals / rate)
        if rc != 0:
            raise RuntimeError(f"batched Poisson run failed: status {rc} ({eng.last_error()})")
        health = eng.counters()
    if health["late_publish"] or health["retried"]:
        # a result that reached the host late or had to be run again is a
        # defect of the engine, not noise: the run fails
        raise RuntimeError(f"batch engine health: {health} ({eng.last_error()})")
    batched = [to_permanent(r, q[0]) for r, q in zip(res, reqs)]

    # plain path: one C-ABI call per matrix, back to back (closed loop)
    calls = {"real": lambda a: api.dense_f64(a), "complex": lambda a: api.dense_c128(a),
             "int01": lambda a: api.int_exact(a.astype(np.int64))}
    plain, plain_lat = [], []
    t0 = time.perf_counter()
    for kind, a, _ in reqs:
        t1 = time.perf_counter()
        plain.append(calls[kind](a))
        plain_lat.append(time.perf_counter() - t1)
    plain_s = time.perf_counter() - t0
    plain_rate = count / plain_s

    # equal accuracy: batched vs plain, per request
    worst, mismatched = 0.0, 0
    for (kind, _, _), b, p in zip(reqs, batched, plain):
        if kind == "int01":
            mismatched += int(b.exact != p.exact)
        else:
            worst = max(worst, abs(b.value - p.value) / abs(p.value))
    pl = np.sort(np.array(plain_lat))
    # latency at an equal offered load that the plain path can sustain (80%
    # of its own capacity): open-loop Poisson arrivals through both paths
    m = min(equal_load_count, count)
    rate_eq = load * plain_rate
    arr_eq = unit_arrivals[:m] / rate_eq
    lat_plain = np.sort(plain_replay(reqs[:m], arr_eq, calls))
    with BatchEngine() as eng:
        eng.tuning = tuning
        rc, _, lat_b, st_eq = eng.replay(reqs[:m], arr_eq)
        if rc != 0:
            raise RuntimeError(f"batched equal-load run failed: status {rc} ({eng.last_error()})")
        health_eq = eng.counters()
    if health_eq["late_publish"] or health_eq["retried"]:
        raise RuntimeError(f"batch engine health: {health_eq}")
    equal_load = {"requests": m, "offered_rate": rate_eq,
                  "plain_p50_ms": float(lat_plain[m // 2] * 1e3), "plain_p99_ms": float(lat_plain[int(0.99 * (m - 1))] * 1e3),
                  "batched_p50_ms": st_eq.p50_s * 1e3, "batched_p99_ms": st_eq.p99_s * 1e3,
                  "rule": "the first requests of the same stream, Poisson arrivals at 80% of the plain path's "
                          "capacity, through one call per matrix (open loop) and through the batch engine"}
    work_flops = sum(request_flops(kind, a.shape[0]) for kind, a, _ in reqs)
    bound = None
    if valu_tflops:
        work_s = work_flops / (valu_tflops * 1e12)
        bound = {"speedup_bound": plain_s / work_s, "work_s_at_valu_rate": work_s, "plain_s": plain_s,
                 "valu_tflops": valu_tflops, "work_flops": work_flops,
                 "rule": "plain wall time / (algorithmic FP64 flops of the stream at the dense n=40 kernel's "
                         "measured TFLOP/s): the ratio with no per-call cost at all"}
    return {
        "workload": f"{count} requests, Poisson arrivals, n uniform {n_lo}..{n_hi}, 50% real uniform[0,1) / 30% "
                    f"complex Gaussian / 20% 0-1 density 0.5 (exact), numpy default_rng({seed})",
        "perms_per_s": st.perms_per_s, "offered_rate": rate, "load": load,
        "p50_ms": st.p50_s * 1e3, "p99_ms": st.p99_s * 1e3, "max_ms": st.max_s * 1e3,
        "batches": st.batches, "mean_batch": st.mean_batch, "device_busy": st.device_busy,
        "late_publish": sat.late_publish + st.late_publish, "retried": sat.retried + st.retried,
        "capacity_perms_per_s": sat.perms_per_s, "capacity_batches": sat.batches,
        "plain_perms_per_s": plain_rate, "plain_call_p50_ms": float(pl[len(pl) // 2] * 1e3),
        "plain_call_p99_ms": float(pl[int(0.99 * (len(pl) - 1))] * 1e3),
        "speedup_vs_plain": sat.perms_per_s / plain_rate,
        "max_rel_diff_vs_plain": worst, "int_mismatches_vs_plain": mismatched,
        "tuning": {"fused": tuning.fused, "waves_per_cu": tuning.waves_per_cu, "log2_chunk": tuning.log2_chunk},
        "compute_bound": bound,
        "equal_load": equal_load,
    }


def secondary() -> dict:
    import superman_amd as S
    from superman_amd import workloads as W

    out = {}
    a = W.dense_uniform(24, 24)
    with S.DensePlan(a, "kahan") as p:
        p.run()
        rs = [p.run() for _ in range(3)]
    one = S.dense_f64(a)
    out["dense_uniform_n24"] = {"kernel_ms": float(np.mean([r.kernel_ms for r in rs])), "value": rs[-1].value,
                                "err_est": rs[-1].err_est, "one_call_ms": one.seconds * 1e3}
    b = W.sparse01(40, 40).astype(np.int64)
    with S.CodegenPlan(b) as p:
        p.run()
        rs = [p.run() for _ in range(2)]
        info = p.codegen_info()
        kname = p.info()["kernel"]
    out["sparse01_n40_codegen"] = {"kernel_ms": float(np.mean([r.kernel_ms for r in rs])), "exact": str(rs[-1].exact),
                                   "compile_s": info["compile_seconds"], "kernel": kname}
    with S.DensePlan(b, "exact", int_walk="sparse") as p:
        p.run()
        rs = p.run()
        out["sparse01_n40_sparse_walk"] = {"kernel_ms": rs.kernel_ms, "exact": str(rs.exact),
                                           "kernel": p.info()["kernel"],
                                           "walk": "compressed columns read at run time (no code generation)"}
    d = S.permanent_decomposed(b.astype(np.float64))  # warm (first-call code loading)
    d = S.permanent_decomposed(b.astype(np.float64))
    out["sparse01_n40_decomposed"] = {"seconds": d.seconds, "kernel_ms": d.kernel_ms, "exact": str(d.exact),
                                      "launches": d.launches,
                                      "pipeline": "DM pruning + components + DecomposePerman compressions, "
                                                  "leaves through the exact integer kernels"}
    u = W.dense_uniform(40, 40)
    with S.DensePlan(u, "kahan") as p:
        r64 = p.run()
    with S.DensePlan(u, "f32_product") as p:
        rf = p.run()
    out["dense_uniform_n40_f32_product"] = {
        "kernel_ms": rf.kernel_ms, "value": rf.value, "err_est": rf.err_est,
        "rel_err_est": rf.err_est / abs(rf.value), "fp64_kernel_ms": r64.kernel_ms, "fp64_value": r64.value,
        "fp64_err_est": r64.err_est, "rel_diff_vs_fp64": abs(rf.value - r64.value) / abs(r64.value),
        "variant": "rows scaled by powers of two, FP64 row sums, FP32 products of n-1 factors, compensated FP64 "
                   "accumulation; nonnegative uniform[0,1) n=40, numpy default_rng(40)"}
    z = W.haar_block(32, 32)
    with S.DensePlan(z, "kahan") as p:
        p.run()
        rs = [p.run() for _ in range(3)]
    v = rs[-1].value
    out["haar_complex_n32"] = {"kernel_ms": float(np.mean([r.kernel_ms for r in rs])), "re": v.real, "im": v.imag,
                               "err_est": rs[-1].err_est,
                               "tflops": 8 * 32 * float(1 << 31) / (np.mean([r.kernel_ms for r in rs]) * 1e-3) / 1e12}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="kahan", choices=["kahan", "dd"])
    ap.add_argument("--n", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=20000)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
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
    # the same permanent with chunks 64 times shorter (exact rebuilds at other
    # places, so every factor is rounded differently): a check of the error
    # estimate that shares no products with the timed run
    with S.DensePlan(a, args.precision) as p3:
        L_short = max(2, info["log2_chunk"] - 6)
        p3.set_log2_chunk(L_short)
        r_short = p3.run()

    ms_per_step = dt * 1e3 / args.steps
    flops = 2.0 * n * float(1 << (n - 1))  # algorithmic 2n flops per subset
    k_ms = float(np.mean(kernel_ms))
    achieved = flops / (k_ms * 1e-3) / 1e12
    # Issue ceiling of any Gray-code Ryser under the 2n convention: at least
    # 2n-1 FP64 VALU instructions per subset (n adds, n-2 multiplies, one
    # FMA into the sum); the peak counts 2 flops (an FMA) per instruction
    # slot, so 2n algorithmic flops in 2n-1 slots reach 2n/(2(2n-1)) of it.
    ceiling = PEAK_FP64_TFLOPS * (2.0 * n) / (2.0 * (2.0 * n - 1.0))
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
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": measured_traffic(info["kernel"]),
                     "ceiling": ceiling, "frac_of_ceiling": achieved / ceiling,
                     "ceiling_rule": "2n-1 FP64 VALU instructions per subset (n adds, n-2 multiplies, 1 FMA) at "
                                     "the peak's issue rate: 2n/(2(2n-1)) of peak; >= 60% of peak is out of reach "
                                     "for any Gray-code Ryser under the 2n-flop convention",
                     "kernel": info["kernel"], "kernel_ms": k_ms,
                     "flops_per_launch": flops, "flop_convention": "2n per subset (add, mul, accumulate; FMA = 2)",
                     "counters": measured_counters(info["kernel"], k_ms)},
        "result": {"value": r.value, "err_est": r.err_est, "cond": r.cond,
                   f"{other}_value": r_other.value, f"{other}_kernel_ms": r_other.kernel_ms,
                   "other_chunk_log2": L_short, "other_chunk_value": r_short.value,
                   "other_chunk_rel_diff": abs(r_short.value - r.value) / abs(r.value),
                   "other_chunk_within_estimates": bool(abs(r_short.value - r.value) <= r.err_est + r_short.err_est)
                   if args.precision == "kahan" else None},
        "geometry": info,
        "device": dev,
    }
    if not args.no_secondary:
        out["secondary"] = secondary()
    if not args.no_serving:
        out["serving"] = serving(args.serving_requests, valu_tflops=achieved)
        # small orders, where each plain call is mostly per-call cost
        out["serving_small_n"] = serving(args.serving_requests, seed=2027, n_lo=8, n_hi=20, valu_tflops=achieved)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
