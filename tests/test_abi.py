"""The C-ABI library loads and exports every function include/lqcd.h declares
(no compute call: runs without a GPU)."""
import ctypes
import os
import re

import pytest

from lqcd import _ffi as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "lqcd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lqcd_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared_functions()
    assert "lqcd_hmc_trajectory" in names and "lqcd_metropolis_sweep" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(F.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert sorted(F.SIGNATURES) == declared_functions()


def test_no_device_is_an_error_not_a_fallback():
    # on a machine without a GPU the context call must fail loudly
    lib = F.lib()
    n = ctypes.c_int(-1)
    rc = lib.lqcd_device_count(ctypes.byref(n))
    if rc != F.OK or n.value == 0:
        import lqcd
        with pytest.raises(lqcd.LqcdError):
            lqcd.Context(0)
