"""C ABI checks that need no GPU: the library loads, exports every symbol
include/*.h declares, and rejects bad arguments before touching a device."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from superman_amd import _lib


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(spm_[A-Za-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "spm_dense_f64" in names and "spm_plan_run" in names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_prototype_table_matches_header():
    assert set(_lib.PROTOTYPES) == set(declared_symbols())


def test_strerror_and_version():
    L = _lib.lib()
    assert L.spm_version().decode().startswith("superman_amd")
    for code in range(0, -11, -1):
        assert L.spm_strerror(code).decode() != "unknown status"


def test_invalid_arguments_rejected_without_gpu():
    L = _lib.lib()
    r = _lib.SpmResult()
    a = np.ones((4, 4))
    dp = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert L.spm_dense_f64(dp, 0, 1, 0, None, ctypes.byref(r)) == _lib.SPM_EINVAL
    assert L.spm_dense_f64(dp, 64, 1, 0, None, ctypes.byref(r)) == _lib.SPM_EINVAL
    assert L.spm_dense_f64(None, 4, 1, 0, None, ctypes.byref(r)) == _lib.SPM_EINVAL
    assert L.spm_dense_f64(dp, 4, 7, 0, None, ctypes.byref(r)) == _lib.SPM_EINVAL
    bad = a.copy()
    bad[1, 2] = np.nan
    assert L.spm_dense_f64(bad.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 4, 1, 0, None,
                           ctypes.byref(r)) == _lib.SPM_EINVAL
    h = ctypes.c_void_p()
    assert L.spm_plan_dense_f64(dp, 0, 1, 0, ctypes.byref(h)) == _lib.SPM_EINVAL
    assert L.spm_plan_run(None, None, ctypes.byref(r)) == _lib.SPM_EINVAL


def test_python_api_raises_without_device():
    import superman_amd as S
    if S.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(_lib.SpmError):
        S.permanent(np.ones((3, 3)))


def test_int_walk_selection_rejected_without_gpu():
    L = _lib.lib()
    plan = ctypes.c_void_p()
    a = np.full((6, 6), 9, dtype=np.int64)  # row sums 54 >= 2^3: the FP64 walk does not apply
    ap = a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    assert L.spm_plan_int_walk(ap, 6, _lib.SPM_INT_WALK_FP64, 0, ctypes.byref(plan)) == _lib.SPM_ERANGE
    assert L.spm_plan_int_walk(ap, 6, 7, 0, ctypes.byref(plan)) == _lib.SPM_EINVAL


def test_f32_product_mode_rejects_what_it_cannot_serve():
    # nonnegative real matrices only; checked before any device call
    L = _lib.lib()
    r = _lib.SpmResult()
    a = np.full((5, 5), 0.5)
    a[2, 3] = -1e-300
    dp = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert L.spm_dense_f64(dp, 5, _lib.SPM_PREC_F32_PRODUCT, 0, None, ctypes.byref(r)) == _lib.SPM_EINVAL
    c = np.ones((5, 10))  # complex, interleaved
    cp = c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert L.spm_dense_c128(cp, 5, _lib.SPM_PREC_F32_PRODUCT, 0, None, ctypes.byref(r)) == _lib.SPM_EINVAL
