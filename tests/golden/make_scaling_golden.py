"""Writes tests/golden/scaling_golden.json: the reference's own CPU binary
(oracle/_ref, built from its sources by `make -C oracle ref`) run with
scaling=true on small matrices, where DecomposePerman balances every leaf
(IO::scale, IO/IO.h:109-161) before its permanent and divides the scales
out afterwards (CommonFiles/DecomposePerman.h:97-107).

    python tests/golden/make_scaling_golden.py

Inputs are seeded (superman_amd.workloads), so the fixture holds only the
recipe and the reference's printed values.  Nonnegative inputs only: the
balancing divides by signed line sums otherwise, and the reference's own
scaled result of a signed N(0,1) matrix is off by orders of magnitude."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import refbin  # noqa: E402
from superman_amd import workloads as W  # noqa: E402

CASES = [
    # name, generator, n, seed, scaling_iteration_no, scale_into
    ("uniform_n12", "dense_uniform", 12, 512, 100, 2),
    ("uniform_n16_it5", "dense_uniform", 16, 516, 5, 3),
    ("sparse01_n34", "sparse01", 34, 534, 100, 2),
]


def matrix(gen, n, seed):
    return getattr(W, gen)(n, seed)


def main():
    out = {"cases": []}
    for name, gen, n, seed, it, into in CASES:
        a = matrix(gen, n, seed)
        plain, _, _ = refbin.run(a, threads=1)
        scaled, _, text = refbin.run(a, threads=1, extra=("scaling=true", f"scaling_iteration_no={it}",
                                                          f"scale_into={into}"))
        out["cases"].append({"name": name, "generator": gen, "n": n, "seed": seed, "scaling_iteration_no": it,
                             "scale_into": into, "reference_scaled": scaled, "reference_plain": plain})
        print(name, plain, scaled, flush=True)
    with open(os.path.join(HERE, "scaling_golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
