"""Runner for the reference's own CPU path (oracle/_ref/superman_ref_cpu).

TEST INFRASTRUCTURE ONLY.  The binary is compiled by `make -C oracle ref`
from the reference sources, exists only where those sources do, and is used
to write and cross-check the golden fixtures under tests/golden/.

The binary speaks the reference's own protocol: `pid=1 repo_dir=<d>/
filename=<f.mtx> mode=cpu thread_count=T [complex=true]` (IO/IO.h:163-401);
with pid=1 it posts a one-int "recompile" message into
`<d>/build/wrapper_pipe_c2p_0.msg` (CommonFiles/Helpers.h:71-116), so every
run gets a fresh scratch directory.  The result line is
`Permanent: <value> - Computed in: <s> seconds.` (main.cpp:122-131,
Algorithms/Complex/cpuComputeComplex.h:127-137).
"""
from __future__ import annotations

import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "_ref", "superman_ref_cpu")

_RE_REAL = re.compile(r"Permanent: (\S+) - Computed in: (\S+) seconds")
_RE_CPLX = re.compile(r"Permanent: \(([^,]+),([^)]+)\) - Computed in: (\S+) seconds")


def available() -> bool:
    return os.path.exists(BIN)


def write_mtx(path: str, a: np.ndarray) -> None:
    """Coordinate MTX with 1-based indices and 17 significant digits (round trip)."""
    n = a.shape[0]
    cplx = np.iscomplexobj(a)
    lines = []
    for i in range(n):
        for j in range(n):
            v = a[i, j]
            if cplx:
                lines.append(f"{i + 1} {j + 1} {v.real:.17g} {v.imag:.17g}")
            elif v != 0:
                lines.append(f"{i + 1} {j + 1} {v:.17g}")
    with open(path, "w") as f:
        f.write(f"{n} {n} {len(lines)}\n")
        f.write("\n".join(lines))
        f.write("\n")


def run(a: np.ndarray, threads: int = 1, extra: tuple[str, ...] = (), timeout: float = 3600.0):
    """Run the reference CPU binary; returns (value, seconds, stdout)."""
    if not available():
        raise RuntimeError("reference binary not built (make -C oracle ref)")
    cplx = np.iscomplexobj(a)
    with tempfile.TemporaryDirectory(prefix="refrun_") as d:
        os.makedirs(os.path.join(d, "build"))
        mtx = os.path.join(d, "m.mtx")
        write_mtx(mtx, a)
        args = [BIN, "pid=1", f"repo_dir={d}/", f"filename={mtx}", "mode=cpu",
                f"thread_count={threads}"]
        if cplx:
            args.append("complex=true")
        args += list(extra)
        out = subprocess.run(args, capture_output=True, text=True, timeout=timeout, cwd=d)
        if out.returncode != 0:
            raise RuntimeError(f"reference run failed ({out.returncode}):\n{out.stdout}\n{out.stderr}")
    text = out.stdout
    if cplx:
        m = _RE_CPLX.search(text)
        if not m:
            raise RuntimeError("no result line:\n" + text)
        return complex(float(m.group(1)), float(m.group(2))), float(m.group(3)), text
    m = _RE_REAL.search(text)
    if not m:
        raise RuntimeError("no result line:\n" + text)
    return float(m.group(1)), float(m.group(2)), text
