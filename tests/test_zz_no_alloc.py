"""
saship_kernel_call_noalloc: the evaluation of saship_kernel_call with device
inputs and outputs and the library's fixed scratch instead of allocated
buffers.  Its results must equal saship_kernel_call's on the same buffers,
on consecutive calls that reuse the scratch (the second one smaller), and it
must refuse calls that do not fit the scratch.
"""

import ctypes as ct

import numpy as np
import pytest

from gpu_helpers import hip_model
from sasmodels_amd import _lib, direct_model
from sasmodels_amd.details import make_kernel_args

pytestmark = pytest.mark.gpu

KIND = {"1d": 0, "2d": 1, "mag": 2}


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _noalloc(model, kind, q, det, val, cutoff=0.0, reff=0, start=0, stop=None, result=None):
    """Raw result of saship_kernel_call_noalloc (q, result on the device)."""
    torch = _torch()
    qd = torch.as_tensor(np.ascontiguousarray(q, "d"), device="cuda")
    nq = q.size if kind == 0 else q.size//2
    nout = 2 if kind == 0 and model.info.have_Fq else 1
    buf = np.ascontiguousarray(getattr(det, "buffer", det), np.int32)
    val = np.ascontiguousarray(val, "d")
    stop = int(buf[-4]) if stop is None else stop
    # room for two outputs per q whatever the model writes
    rd = torch.full((2*nq + 4,), np.nan, dtype=torch.float64, device="cuda")
    if result is not None:
        rd[:nout*nq + 4] = torch.as_tensor(np.asarray(result, "d"), device="cuda")
    torch.cuda.synchronize()
    _lib.check(_lib.lib().saship_kernel_call_noalloc(
        model._model_handle(), kind, nq, ct.c_void_p(qd.data_ptr()), buf.ctypes.data, buf.size,
        val.ctypes.data, val.size, float(cutoff), int(reff), int(start), int(stop),
        ct.c_void_p(rd.data_ptr()), ct.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    return rd.cpu().numpy()[:nout*nq + 4]


def _case(name, quad, q_vectors, pars):
    model = hip_model(name, quad)
    k = model.make_kernel(q_vectors)
    mesh = direct_model.get_mesh(model.info, dict(pars), dim=k.dim)
    det, val, mag = make_kernel_args(k, mesh)
    return model, k, det, val, mag


def _compare(name, quad, q_vectors, pars, cutoff=0.0, reff=0, separable=True):
    model, k, det, val, mag = _case(name, quad, q_vectors, pars)
    kind = 0 if k.dim == "1d" else (2 if mag else 1)
    q = q_vectors[0] if kind == 0 else np.column_stack([np.ravel(v) for v in q_vectors]).ravel()
    k.set_separable(separable)
    k._call_kernel(det, val, cutoff, mag, reff)
    want = k.result.copy()
    got = _noalloc(model, kind, q, det, val, cutoff, reff)
    k.release()
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-300)
    return got


def _grid(n, lim=0.3):
    x = np.linspace(-lim, lim, n)
    return [a.ravel() for a in np.meshgrid(x, x)]


def test_noalloc_matches_kernel_call_twice():
    """Two calls in a row on the shared scratch, the second with fewer q and
    mesh points, each equal to saship_kernel_call on the same inputs."""
    big = dict(radius=20, length=400, radius_pd=0.1, radius_pd_n=35, radius_pd_type="lognormal",
               length_pd=0.1, length_pd_n=35, length_pd_type="lognormal", sld=4, sld_solvent=1)
    small = dict(big, radius_pd_n=9, length_pd_n=7, radius=31)
    for sep in (True,):
        _compare("cylinder", 76, [np.logspace(-3, 0, 1000)], big)
        _compare("cylinder", 76, [np.logspace(-3, 0, 200)], small)
    # the adaptive rule: here the generic mesh kernel (the separable form of
    # this rule runs through a batch's buffers), compared with the kernel
    # call forced onto the same path
    _compare("cylinder", 0, [np.logspace(-3, 0, 700)], dict(big, radius_pd_n=15, length_pd_n=11), separable=False)
    _compare("cylinder", 0, [np.logspace(-2, 0, 90)], dict(big, radius_pd_n=5, length_pd_n=3), separable=False)


def test_noalloc_2d_magnetic_structure():
    qx, qy = _grid(64)
    _compare("ellipsoid", 0, [qx, qy], dict(radius_polar=20, radius_equatorial=400, theta=60, phi=60,
                                            theta_pd=10, theta_pd_n=8, phi_pd=20, phi_pd_n=8,
                                            radius_equatorial_pd=0.1, radius_equatorial_pd_n=6))
    qx, qy = _grid(32)
    _compare("core_shell_sphere", 0, [qx, qy], dict(radius=60, thickness=10, sld_core_M0=2, sld_shell_M0=1,
                                                    sld_solvent_M0=0.5, sld_core_mtheta=30, sld_core_mphi=40,
                                                    up_frac_i=0.3, up_frac_f=0.8, up_theta=45,
                                                    radius_pd=0.1, radius_pd_n=5))
    _compare("hardsphere", 0, [np.logspace(-3, 0, 300)], dict(radius_effective=50, volfraction=0.2))
    _compare("sphere", 0, [np.logspace(-3, 0, 77)], dict(radius=50, radius_pd=0.1, radius_pd_n=35),
             cutoff=1e-5, reff=1)


def test_noalloc_chunked_accumulates():
    """pd_start > 0 adds to the device result, as saship_kernel_call adds to the host one."""
    model, k, det, val, mag = _case("sphere", 0, [np.logspace(-3, 0, 200)],
                                    dict(radius=50, radius_pd=0.1, radius_pd_n=35))
    q = np.logspace(-3, 0, 200)
    k._call_kernel(det, val, 0.0, False, 0)
    whole = k.result.copy()
    # the same two chunks through saship_kernel_call
    buf = np.ascontiguousarray(det.buffer, np.int32)
    vals = np.ascontiguousarray(val, "d")
    want = np.zeros_like(whole)
    for start, stop in ((0, 13), (13, 35)):
        _lib.check(_lib.lib().saship_kernel_call(k._handle, 0, buf.ctypes.data, buf.size, vals.ctypes.data,
                                                 vals.size, 0.0, 0, start, stop, want.ctypes.data))
    k.release()
    part = _noalloc(model, 0, q, det, val, start=0, stop=13)
    part = _noalloc(model, 0, q, det, val, start=13, stop=35, result=part)
    np.testing.assert_allclose(part, want, rtol=1e-12, atol=1e-300)
    # chunking changes only the FP64 summation order
    np.testing.assert_allclose(part, whole, rtol=1e-12, atol=1e-300)


def test_noalloc_rejects_what_does_not_fit():
    torch = _torch()
    model = hip_model("sphere", 0)
    nq = 5_000_000                    # 20 MB of FP32 q plus 80 MB of partial sums: over the scratch
    qd = torch.full((nq,), 0.1, dtype=torch.float64, device="cuda")
    rd = torch.zeros(2*nq + 4, dtype=torch.float64, device="cuda")
    k = model.make_kernel([np.array([0.1])])
    det, val, _ = make_kernel_args(k, direct_model.get_mesh(model.info, dict(radius=50), dim="1d"))
    k.release()
    buf = np.ascontiguousarray(det.buffer, np.int32)
    val = np.ascontiguousarray(val, "d")
    rc = _lib.lib().saship_kernel_call_noalloc(model._model_handle(), 0, nq, ct.c_void_p(qd.data_ptr()),
                                                buf.ctypes.data, buf.size, val.ctypes.data, val.size, 0.0, 0,
                                                0, 1, ct.c_void_p(rd.data_ptr()), None)
    assert rc == -1
    assert b"scratch" in _lib.lib().saship_last_error()
    host = np.zeros(2*10 + 4)
    rc = _lib.lib().saship_kernel_call_noalloc(model._model_handle(), 0, 10, ct.c_void_p(qd.data_ptr()),
                                                buf.ctypes.data, buf.size, val.ctypes.data, val.size, 0.0, 0,
                                                0, 1, host.ctypes.data, None)
    assert rc == -1
    assert b"device memory" in _lib.lib().saship_last_error()
