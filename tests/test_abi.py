"""
The C ABI without a GPU: libsaship.so loads, exports every function
include/saship.h declares, and rejects bad arguments with a message before
touching a device.
"""

import ctypes as ct
import os
import re

import pytest

from sasmodels_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "saship.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(saship_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = _lib.lib()
    declared = _declared()
    assert "saship_kernel_call" in declared
    missing = [name for name in declared if not hasattr(lib, name)]
    assert not missing, missing
    assert set(declared) == set(_lib.SIGNATURES), "ctypes signature table out of date"


def test_error_reporting_without_device():
    lib = _lib.lib()
    out = ct.c_void_p()
    assert lib.saship_model_create(7, 0, ct.byref(out)) == -1
    assert b"unknown model id" in lib.saship_last_error()
    assert lib.saship_model_create(1, -5, ct.byref(out)) == -1
    assert lib.saship_kernel_call(None, 0, None, 0, None, 0, 0.0, 0, 0, 0, None) == -1
    assert b"NULL" in lib.saship_last_error()
    with pytest.raises(_lib.HipError):
        _lib.check(lib.saship_kernel_release(None) - 1)
    n = _lib.device_count()
    assert n >= 0
