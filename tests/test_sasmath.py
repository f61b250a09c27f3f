"""
Accuracy of the FP32 special functions of the kernels (sasmodels_amd/csrc/
sasmath.h), compiled for the host, against scipy in FP64 over the argument
range of the benchmark configurations, 1e-3 .. 3e3.  The error is measured
relative to the function's envelope (its size away from the zeros:
min(1, 1/x) for sin(x)/x, min(1, 3/x^2) for 3 j1(x)/x, min(1, sqrt(8/pi) x^-3/2)
for 2 J1(x)/x), since near a zero no FP32 evaluation of an FP32 argument has
a small relative error.  For the *_f functions the argument is first rounded
to FP32 and the reference is evaluated at that rounded argument.
"""

import ctypes as ct
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy import special

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("sasmath") / "libsasmath_host.so")
    subprocess.run([hipcc, "-O2", "-fPIC", "-shared", "-o", out, os.path.join(HERE, "sasmath_host.cpp")],
                   check=True, capture_output=True)
    L = ct.CDLL(out)
    for name in ("t_sin_f", "t_sinc_f", "t_sph_j1c_f", "t_j1c_f", "t_sinc_d", "t_sph_j1c_d", "t_j1c_d",
                 "t_sph_j1c_dd"):
        getattr(L, name).restype = ct.c_double
        getattr(L, name).argtypes = [ct.c_double]
    return L


X = np.concatenate([np.logspace(-3, np.log10(3e3), 40000), np.linspace(0.5, 30, 20000)])


def _sinc(x):
    return np.sin(x)/x


def _sph(x):
    return 3*special.spherical_jn(1, x)/x


def _j1c(x):
    return 2*special.j1(x)/x


def _env(kind, x):
    if kind == "sinc":
        return np.minimum(1.0, 1.0/x)
    if kind == "sph":
        return np.minimum(1.0, 3.0/x**2)
    return np.minimum(1.0, np.sqrt(8/np.pi)*x**-1.5)


@pytest.mark.parametrize("name,ref,kind,bound", [
    ("t_sinc_f", _sinc, "sinc", 7e-7), ("t_sph_j1c_f", _sph, "sph", 7e-7), ("t_j1c_f", _j1c, "j1c", 7e-7),
    ("t_sinc_d", _sinc, "sinc", 7e-7), ("t_sph_j1c_d", _sph, "sph", 7e-7), ("t_j1c_d", _j1c, "j1c", 7e-7),
])
def test_envelope_error(lib, name, ref, kind, bound):
    f = np.vectorize(getattr(lib, name))
    x = X.astype(np.float32).astype(np.float64) if name.endswith("_f") else X
    got = f(x)
    err = np.abs(got - ref(x))/_env(kind, x)
    print(name, "max envelope error %.3g at x = %.4g" % (err.max(), x[np.argmax(err)]))
    assert err.max() <= bound


def test_sin_reduction_large_arguments(lib):
    """sin of FP32 arguments up to 1e5 (the FP64 reduction above 1e3): within
    a few FP32 ulps of 1 absolute, whatever the size of the argument."""
    x = np.logspace(0, 5, 20001).astype(np.float32).astype(np.float64)
    got = np.vectorize(lib.t_sin_f)(x)
    assert np.max(np.abs(got - np.sin(x))) <= 3e-7


def test_fp64_series_of_3j1x_x(lib):
    """The FP64 3 j1(x)/x of the FP64 models, relative to its envelope: the
    closed form just above the series' switch at x = 0.25 cancels about six
    bits (sin x - x cos x ~ x^3/3)."""
    x = np.logspace(-4, 2, 5001)
    got = np.vectorize(lib.t_sph_j1c_dd)(x)
    err = np.abs(got - _sph(x))/_env("sph", x)
    assert err.max() <= 1e-14
