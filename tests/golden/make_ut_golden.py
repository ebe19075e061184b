"""Writes the UT-order golden fixture (run where the reference is mounted).

    python tests/golden/make_ut_golden.py

The reference ships one generated SpaRyser kernel,
Algorithms/GPU/CodeGeneration/generatedKernels.cuh, produced from
ExampleMatrices/bcsstk01.mtx read as undirected (IO/IO.h:403-515 with
undirected=true) and ordered by IO::UTOrder (IO/IO.h:926-1002).  Its
c{j}included functions list, for every column j of the ordered matrix, the
rows it updates and the value it adds (rows below the register area k = 15
as `reg<i>`, the others as `x[totalThreadCount * (i - k) + ...]`).  This
script stores, as data only:
  bcsstk01_undirected.npy       the input matrix (from the .mtx file)
  ut_bcsstk01_reference.json    {column: [[row, printed value], ...]} parsed
                                from that generated kernel.
"""
import json
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
K_REGISTER_AREA = 15  # `reg0..reg14` in the generated file


def read_mtx_undirected(path):
    lines = [ln for ln in open(path) if not ln.startswith("%")]
    n, m, nl = map(int, lines[0].split())
    a = np.zeros((n, n))
    for ln in lines[1:1 + nl]:
        i, j, v = ln.split()[:3]
        i, j, v = int(i) - 1, int(j) - 1, float(v)
        if v == 0:
            continue
        a[i, j] = v
        if i != j:
            a[j, i] = v
    return a


def main():
    a = read_mtx_undirected(os.path.join(REF, "ExampleMatrices", "bcsstk01.mtx"))
    np.save(os.path.join(HERE, "bcsstk01_undirected.npy"), a, allow_pickle=False)
    text = open(os.path.join(REF, "Algorithms", "GPU", "CodeGeneration", "generatedKernels.cuh")).read()
    cols = {}
    for m in re.finditer(r"void c(\d+)included\((.*?)\n\}", text, re.S):
        j = int(m.group(1))
        body = m.group(2)
        ent = [[int(r), v] for r, v in re.findall(r"reg(\d+) \+= (-?[\d.]+);", body)]
        ent += [[int(r) + K_REGISTER_AREA, v]
                for r, v in re.findall(r"x\[totalThreadCount \* (\d+) \+ globalThreadID\] \+= (-?[\d.]+);", body)]
        cols[str(j)] = sorted(ent)
    with open(os.path.join(HERE, "ut_bcsstk01_reference.json"), "w") as f:
        json.dump({"source": "Algorithms/GPU/CodeGeneration/generatedKernels.cuh (c{j}included)",
                   "input": "ExampleMatrices/bcsstk01.mtx, undirected", "n": int(a.shape[0]),
                   "value_format": "std::to_string(double), 6 decimals", "columns": cols}, f, indent=0)
    print("columns:", len(cols))


if __name__ == "__main__":
    main()
