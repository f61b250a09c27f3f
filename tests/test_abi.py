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
    assert lib.saship_model_create(99, 0, ct.byref(out)) == -1
    assert b"unknown model id" in lib.saship_last_error()
    assert lib.saship_model_create(1, -5, ct.byref(out)) == -1
    assert lib.saship_kernel_call(None, 0, None, 0, None, 0, 0.0, 0, 0, 0, None) == -1
    assert b"NULL" in lib.saship_last_error()
    with pytest.raises(_lib.HipError):
        _lib.check(lib.saship_kernel_release(None) - 1)
    n = _lib.device_count()
    assert n >= 0


def test_one_hip_runtime_per_process():
    """Loading libsaship.so and then torch leaves a single HIP runtime mapped
    (two runtimes would hide the device from torch and break stream sharing)."""
    import subprocess
    import sys
    code = (
        "import re\n"
        "from sasmodels_amd import _lib\n"
        "_lib.lib()\n"
        "import torch\n"
        "maps = open('/proc/self/maps').read()\n"
        "print(len(set(re.findall(r'\\S*libamdhip64\\S*', maps))))\n"
    )
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().splitlines()[-1] == "1"
