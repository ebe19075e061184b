import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (run with -m gpu)")
    config.addinivalue_line("markers", "slow: long-running case")


def load_golden():
    with open(os.path.join(GOLDEN_DIR, "golden.json")) as f:
        g = json.load(f)
    return g["cases"]


def golden_matrix(case) -> np.ndarray:
    return np.load(os.path.join(GOLDEN_DIR, case["file"]), allow_pickle=False)


@pytest.fixture(scope="session")
def golden():
    return {c["name"]: c for c in load_golden()}


def rel_err(x, ref) -> float:
    ref = complex(ref)
    d = abs(complex(x) - ref)
    return d / abs(ref) if ref != 0 else d
