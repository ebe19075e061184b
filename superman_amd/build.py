"""Builds libsuperman_amd.so in-tree with hipcc for gfx950.

    python -m superman_amd.build [-j JOBS] [--verbose]

Every HIP translation unit is compiled with
    -Xclang -target-feature -Xclang -packed-fp32-ops
(no packed FP32 instructions are generated), objects go to
superman_amd/_build/, and the shared library is linked next to this file.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shlex
import subprocess
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libsuperman_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# (source, extra flags).  .hip files carry device code; .cpp files are host
# code compiled by the same driver (HIP runtime API).
SOURCES = [  # the longest compiles first, so that the job pool starts them early
    ("int_sparse_26.hip", []),  # the sparse kernel's orders, split to compile in parallel
    ("int_sparse_37.hip", []),
    ("int_sparse_32.hip", []),
    ("int_sparse_42.hip", []),
    ("int_sparse_57.hip", []),
    ("int_sparse_49.hip", []),
    ("int_sparse_46.hip", []),
    ("dense_c128.hip", []),
    ("int_sparse_17.hip", []),
    ("batch.hip", []),
    ("int_dense.hip", []),
    ("dense_f64.hip", []),
    ("int_sparse.hip", []),
    ("int_small.hip", []),
    ("batch_engine.cpp", []),
    ("decompose.cpp", []),
    ("codegen.cpp", []),
    ("capi.cpp", []),
    ("prep.cpp", []),
    ("dense_noalloc.hip", []),
]
LINK_LIBS = ["-lhiprtc", "-lpthread"]


def _compile(src: str, flags: list[str], verbose: bool) -> tuple[str, float, str]:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + flags + ["-c", path, "-o", obj]
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    t = time.time()
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{out.stdout}\n{out.stderr}")
    return obj, time.time() - t, out.stderr


def _headers_mtime() -> float:
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    inc = os.path.join(ROOT, "include")
    for d, _, fs in os.walk(inc):
        hs += [os.path.join(d, f) for f in fs if f.endswith(".h")]
    return max((os.path.getmtime(h) for h in hs), default=0.0)


def _fresh(src: str, defines) -> bool:
    """An object newer than its source and every header (incremental
    builds of the default configuration only)."""
    if defines:
        return False
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if not os.path.exists(obj):
        return False
    t = os.path.getmtime(obj)
    return t > os.path.getmtime(os.path.join(CSRC, src)) and t > _headers_mtime()


def build(jobs: int = 8, verbose: bool = False, defines=(), out: str | None = None,
          build_dir: str | None = None, incremental: bool = False) -> str:
    """Compile and link.  `defines` (NAME=VALUE) and `out`/`build_dir` let a
    tuning sweep build variants of the library away from the package;
    `incremental` recompiles only sources older than their object."""
    global BUILD
    lib_path = out or LIB
    if build_dir:
        BUILD = build_dir
    os.makedirs(BUILD, exist_ok=True)
    extra = ["-D" + d for d in defines]
    t0 = time.time()
    objs = []
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = []
        for s, f in SOURCES:
            if incremental and _fresh(s, defines):
                objs.append(os.path.join(BUILD, os.path.splitext(s)[0] + ".o"))
            else:
                futs.append(ex.submit(_compile, s, f + extra, verbose))
        for fu in futs:
            obj, dt, _ = fu.result()
            objs.append(obj)
            if verbose:
                print(f"  {os.path.basename(obj)}: {dt:.1f}s", flush=True)
    tmp = lib_path + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + LINK_LIBS
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"link failed\n{out.stdout}\n{out.stderr}")
    os.replace(tmp, lib_path)
    if out_is_default(lib_path):
        build_cli(verbose)
    if verbose:
        print(f"built {lib_path} in {time.time() - t0:.1f}s", flush=True)
    return lib_path


CLI = os.path.join(PKG, "superman_gpu")


def out_is_default(lib_path: str) -> bool:
    return os.path.abspath(lib_path) == os.path.abspath(LIB)


def build_cli(verbose: bool = False) -> str:
    """The command-line driver (csrc/cli_main.cpp): host C++ only, linked to
    the in-tree library with an $ORIGIN run path."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
           os.path.join(CSRC, "cli_main.cpp"), "-L" + PKG, "-lsuperman_amd", "-Wl,-rpath,$ORIGIN", "-o", CLI]
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"CLI build failed\n{out.stdout}\n{out.stderr}")
    return CLI


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--verbose", "-v", action="store_true")
    ap.add_argument("-D", dest="defines", action="append", default=[], help="NAME=VALUE for a tuning variant")
    ap.add_argument("--out", help="library path (default: in the package)")
    ap.add_argument("--build-dir", help="object directory (default: superman_amd/_build)")
    ap.add_argument("--incremental", action="store_true", help="recompile only sources newer than their object")
    a = ap.parse_args(argv)
    build(a.jobs, a.verbose, a.defines, a.out, a.build_dir, a.incremental)


if __name__ == "__main__":
    sys.exit(main())
