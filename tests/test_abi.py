"""The C-ABI library loads and exports every symbol include/plasticity_amd.h declares (no GPU calls)."""
import os
import re

import plasticity_amd.native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "plasticity_amd.h")).read()
    return sorted(set(re.findall(r"\b(pa_[a-z_]+)\s*\(", src)))


def test_library_exports_all_declared_symbols():
    L = nat.lib()
    names = declared()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), name
    assert sorted(nat.EXPORTS) == names


def test_error_path_without_gpu_is_reported():
    import ctypes
    import plasticity_amd as pa
    mat = nat.PaMaterial()
    mat.n_slip = 0  # invalid: rejected before any device call
    pb = nat.PaProblem()
    h = ctypes.c_void_p()
    rc = nat.lib().pa_create(ctypes.byref(mat), ctypes.byref(pb), ctypes.byref(h))
    assert rc == 1
    assert b"n_slip" in nat.lib().pa_last_error()
