"""Runs superman_amd/superman_gpu (the command-line driver) on matrices
written to a temporary MatrixMarket file."""
import os
import re
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "superman_amd", "superman_gpu")


def binary():
    if not os.path.exists(BIN):
        from superman_amd import build as B
        B.build_cli()
    return BIN


def write_mtx(path, a, pattern=False, symmetric=False):
    """MatrixMarket coordinate file; `symmetric` writes the lower triangle
    under a symmetric banner (read back with undirected=true)."""
    n = a.shape[0]
    cplx = np.iscomplexobj(a)
    rows = []
    for i in range(n):
        for j in range(n):
            v = a[i, j]
            if v == 0 or (symmetric and j > i):
                continue
            if pattern:
                rows.append(f"{i + 1} {j + 1}")
            elif cplx:
                rows.append(f"{i + 1} {j + 1} {v.real:.17g} {v.imag:.17g}")
            else:
                rows.append(f"{i + 1} {j + 1} {v:.17g}")
    kind = "pattern" if pattern else ("complex" if cplx else "real")
    shape = "symmetric" if symmetric else "general"
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {kind} {shape}\n% written by the tests\n{n} {n} {len(rows)}\n" +
                "\n".join(rows) + "\n")


def run(a=None, *args, pattern=False, symmetric=False, path=None):
    with tempfile.TemporaryDirectory(prefix="spm_cli_") as d:
        if path is None:
            path = os.path.join(d, "m.mtx")
            write_mtx(path, a, pattern, symmetric)
        p = subprocess.run([binary(), f"filename={path}", *args], capture_output=True, text=True, timeout=600)
    return p.returncode, p.stdout, p.stderr


def value(stdout):
    m = re.search(r"Permanent: \(([^,]+),([^)]+)\) - Computed in", stdout)
    if m:
        return complex(float(m.group(1)), float(m.group(2)))
    m = re.search(r"Permanent: (\S+) - Computed in", stdout)
    return float(m.group(1))
