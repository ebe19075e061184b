"""Writes the decomposition fixtures (run where the reference is mounted and
oracle/_ref is built: `make -C oracle ref`).

    python tests/golden/make_decomp_golden.py

For each input the reference's own CPU binary (oracle/_ref/superman_ref_cpu,
mode=cpu) prints, after its DM pruning, component split and DecomposePerman
compressions (main.cpp:58-131, CommonFiles/DecomposePerman.h:76-117):
    Submatrix count: <leaves> - Average N: <n> - Average NNZ: <nnz>
    Permanent: <value> - Computed in: <s> seconds.
Stored as data: decomp_golden.json {name: {file, leaves, avg_n, avg_nnz,
value}} and the inputs that are not already fixtures (decomp_*.npy).
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import refbin  # noqa: E402

_RE = re.compile(r"Submatrix count: (\d+) - Average N: (\S+) - Average NNZ: (\S+)")


def sparse_ints(n, seed, per_row=4, vmax=3):
    """A seeded permutation plus per_row - 1 more random entries per row,
    values in 1..vmax (the config-3 generator with small integer values)."""
    rng = np.random.default_rng(seed)
    p = rng.permutation(n)
    a = np.zeros((n, n))
    for i in range(n):
        a[i, p[i]] = rng.integers(1, vmax + 1)
        others = [c for c in range(n) if c != p[i]]
        for c in rng.choice(others, per_row - 1, replace=False):
            a[i, c] = rng.integers(1, vmax + 1)
    return a


def main():
    cases = {}
    inputs = [(f"c3_sparse01_n40_s{s}", f"c3_sparse01_n40_s{s}.npy", None) for s in range(40, 45)]
    for n, seed, per_row in ((36, 101, 4), (38, 102, 3), (42, 103, 4), (34, 104, 5)):
        name = f"decomp_ints_n{n}_s{seed}"
        inputs.append((name, name + ".npy", sparse_ints(n, seed, per_row)))
    for name, fname, a in inputs:
        path = os.path.join(HERE, fname)
        if a is None:
            a = np.load(path, allow_pickle=False)
        else:
            np.save(path, a, allow_pickle=False)
        value, secs, text = refbin.run(a, threads=8)
        m = _RE.search(text)
        if not m:
            raise RuntimeError("no submatrix line for " + name)
        cases[name] = {"file": fname, "n": int(a.shape[0]), "leaves": int(m.group(1)), "avg_n": float(m.group(2)),
                       "avg_nnz": float(m.group(3)), "value": value, "seconds": secs}
        print(name, cases[name], flush=True)
    with open(os.path.join(HERE, "decomp_golden.json"), "w") as f:
        json.dump({"source": "reference CPU binary (mode=cpu, thread_count=8): Submatrix count / Average N / "
                             "Average NNZ line and the permanent", "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
