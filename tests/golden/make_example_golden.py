"""Writes fixtures from the reference's example matrices (run where the
reference is mounted and oracle/_ref is built: `make -C oracle ref`).

    python tests/golden/make_example_golden.py

Each ExampleMatrices/<name>.mtx is read here with the settings a user of
the reference gives it (binary=true for a `pattern` banner, undirected=true
for a `symmetric` one: IO/IO.h:403-470 reads neither banner word itself),
saved as example_<name>.npy, and run through the reference's own CPU binary
(oracle/_ref/superman_ref_cpu, mode=cpu), whose "Submatrix count / Average N /
Average NNZ" line and permanent go to example_golden.json, with (for
non-0/1 inputs) the double-double permanent of the leaves: d_ss is so
ill-conditioned that FP64 results, the reference's included, differ from
it in the second or third digit.  Matrices the
reference needs more than `--timeout` seconds for are left out (sparse,
curtis54, mycielskian6, mesh1e1 and bcsstk01 each ran over 3 minutes on
8 threads).
"""
import argparse
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402
from oracle import refbin  # noqa: E402
from superman_amd import api  # noqa: E402

EXAMPLES = "/root/reference/ExampleMatrices"
_RE = re.compile(r"Submatrix count: (\d+) - Average N: (\S+) - Average NNZ: (\S+)")


def read_example(path):
    """Coordinate MatrixMarket, 1-based; returns (matrix, pattern, symmetric)."""
    lines = open(path).read().splitlines()
    banner = lines[0].lower() if lines and lines[0].startswith("%%") else ""
    body = [ln for ln in lines if ln.strip() and not ln.startswith("%")]
    n, m, k = (int(t) for t in body[0].split())
    pattern, sym = "pattern" in banner, "symmetric" in banner
    a = np.zeros((n, m))
    for ln in body[1:1 + k]:
        t = ln.split()
        i, j = int(t[0]) - 1, int(t[1]) - 1
        v = 1.0 if pattern else float(t[2])
        if v == 0.0:
            continue
        a[i, j] = v
        if sym and i != j:
            a[j, i] = v
    return a, pattern, sym


def leaves_dd(a):
    """prod over components of the sum of the leaves' double-double
    permanents (the leaves as this library's preprocessing writes them,
    which the leaf statistics above tie to the reference's)."""
    d = api.decompose(a)
    comp = {}
    for c, m in d.leaves:
        hi, lo = O.dense_dd(m, threads=8)
        comp[c] = comp.get(c, 0.0) + (hi + lo)
    v = 1.0
    for x in comp.values():
        v *= x
    return 0.0 if d.structurally_zero else v


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--timeout", type=float, default=240.0)
    ap.add_argument("names", nargs="*", default=["bcspwr02", "d_ss"])
    args = ap.parse_args()
    out = os.path.join(HERE, "example_golden.json")
    cases = json.load(open(out))["cases"] if os.path.exists(out) else {}
    for name in args.names:
        a, pattern, sym = read_example(os.path.join(EXAMPLES, name + ".mtx"))
        try:
            value, secs, text = refbin.run(a, threads=8, timeout=args.timeout)
        except Exception as e:  # too slow for a fixture
            print(name, "skipped:", type(e).__name__, flush=True)
            continue
        m = _RE.search(text)
        fname = f"example_{name}.npy"
        np.save(os.path.join(HERE, fname), a, allow_pickle=False)
        cases[name] = {"file": fname, "n": int(a.shape[0]), "binary": pattern, "undirected": sym,
                       "leaves": int(m.group(1)), "avg_n": float(m.group(2)), "avg_nnz": float(m.group(3)),
                       "value": value, "seconds": secs}
        if not pattern:
            cases[name]["leaves_dd"] = leaves_dd(a)
        print(name, cases[name], flush=True)
    with open(out, "w") as f:
        json.dump({"source": "reference CPU binary (mode=cpu, thread_count=8) on ExampleMatrices/*.mtx",
                   "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
