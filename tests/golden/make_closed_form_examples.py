"""Writes the reference's three closed-form example matrices as fixtures
(run where the reference is mounted):

    python tests/golden/make_closed_form_examples.py

ExampleMatrices/dense.mtx (45 x 45, every entry 0.91), complex.mtx (40 x 40,
every entry i) and sparse.mtx (45 x 45, 483 entries of 0.91) are read as the
reference reads a coordinate file (1-based, IO/IO.h:403-470 and :517-596)
and saved as example_<name>.npy.  Their permanents have closed forms:
dense = 45! * 0.91^45 and complex = 40! * i^40 exactly (with 0.91 the
binary64 number the file's text rounds to, the power taken in exact rational
arithmetic); sparse = 0.91^45 * perm(pattern), whose integer factor the GPU
tests take from the exact integer path.  The closed-form values go to
closed_form_examples.json as decimal strings.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLES = "/root/reference/ExampleMatrices"


def read_coo(path, cplx):
    body = [ln for ln in open(path).read().splitlines() if ln.strip() and not ln.startswith("%")]
    n, m, k = (int(t) for t in body[0].split())
    a = np.zeros((n, m), dtype=np.complex128 if cplx else np.float64)
    for ln in body[1:1 + k]:
        t = ln.split()
        i, j = int(t[0]) - 1, int(t[1]) - 1
        a[i, j] = complex(float(t[2]), float(t[3])) if cplx else float(t[2])
    return a, k


def main():
    out = {}
    for name, cplx in (("dense", False), ("complex", True), ("sparse", False)):
        a, k = read_coo(os.path.join(EXAMPLES, name + ".mtx"), cplx)
        fname = f"example_{name}.npy"
        np.save(os.path.join(HERE, fname), a, allow_pickle=False)
        vals = sorted({complex(v) for v in a.ravel() if v != 0}, key=lambda z: (z.real, z.imag))
        out[name] = {"file": fname, "n": int(a.shape[0]), "nnz": int(k), "distinct_nonzeros": [str(v) for v in vals]}
    c = Fraction(0.91)  # the binary64 value of the file's 0.91
    dense = math.factorial(45) * c ** 45
    out["dense"]["perm"] = float(dense)
    out["complex"]["perm"] = str(math.factorial(40) * (1 if 40 % 4 == 0 else -1))  # 40! i^40 = 40!
    out["sparse"]["entry_power"] = float(c ** 45)
    with open(os.path.join(HERE, "closed_form_examples.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1)[:2000])


if __name__ == "__main__":
    main()
