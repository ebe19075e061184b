"""Builds and runs tests/cpp/dropin_main.cpp (the C++ drop-in headers over
libsuperman_amd.so) with g++ into a temporary directory."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "superman_amd")
_BIN = None


def binary():
    global _BIN
    if _BIN is None:
        out = os.path.join(tempfile.mkdtemp(prefix="spm_dropin_"), "dropin")
        subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp"), "-L", LIBDIR,
                        "-lsuperman_amd", "-Wl,-rpath," + LIBDIR, "-o", out], check=True)
        _BIN = out
    return _BIN


def run(mode, a, algorithm=-1, precision=1):
    a = np.asarray(a)
    n = a.shape[0]
    if mode == "cplx":
        body = " ".join("%r %r" % (float(z.real), float(z.imag)) for z in a.reshape(-1))
    else:
        body = " ".join(repr(float(v)) for v in a.reshape(-1))
    p = subprocess.run([binary()], input="%s %d %d %d %s\n" % (mode, algorithm, precision, n, body),
                       capture_output=True, text=True, timeout=300)
    return p.returncode, p.stdout.strip().splitlines()
