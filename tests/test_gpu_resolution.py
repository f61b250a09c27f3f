"""
Resolution smearing on the device after the batched kernel: against the
host application of the same weight matrices to the unsmeared batch output
(FP64 on both sides) and against the FP64 oracle smeared on the host.
"""

import numpy as np
import pytest

from gpu_helpers import RTOL, hip_model, masked_relerr
from oracle.oracle import OracleKernel
from sasmodels_amd import direct_model, resolution as R
from sasmodels_amd.details import make_kernel_args
from sasmodels_amd.kernelhip import HipBatch

pytestmark = pytest.mark.gpu


def _sets(hm, n, seed):
    rng = np.random.default_rng(seed)
    kl = type("K", (), {"info": hm.info, "dtype": np.dtype("d")})()
    res, calls = [], []
    for s in range(n):
        q = np.logspace(-3, np.log10(0.5), int(rng.integers(20, 300)))
        kind = s % 3
        if kind == 0:
            res.append(R.Pinhole1D(q, rng.uniform(0.02, 0.12)*q))
        elif kind == 1:
            res.append(R.Slit1D(q, q_length=rng.uniform(0.005, 0.05), q_width=0.0))
        else:
            res.append(R.Perfect1D(q))
        pars = dict(radius=rng.uniform(10, 80), length=rng.uniform(50, 600), radius_pd=0.1, radius_pd_n=9,
                    length_pd=0.1, length_pd_n=7)
        calls.append(make_kernel_args(kl, direct_model.get_mesh(hm.info, pars, dim="1d"))[:2])
    return res, calls


@pytest.mark.parametrize("graph", [False, True])
def test_batch_smearing(graph):
    hm = hip_model("cylinder", 76)
    res, calls = _sets(hm, 9, 31)
    batch = HipBatch(hm, [[r.q_calc] for r in res], use_graph=graph)
    plain = batch(calls)
    batch.set_resolution(res)
    for rep in range(2 if graph else 1):
        print("smearing graph", graph, "rep", rep, flush=True)
        smeared = batch(calls)
        for r, p, sm, (det, val) in zip(res, plain, smeared, calls):
            want = r.apply(p)
            assert sm.shape == want.shape
            np.testing.assert_allclose(sm, want, rtol=1e-12, atol=0)
    # against the FP64 oracle smeared on the host
    for r, sm, (det, val) in zip(res, smeared, calls):
        ref = r.apply(OracleKernel(hm.info, [r.q_calc])(det, val, 0.0, False))
        assert masked_relerr(sm, ref) <= RTOL
    batch.set_resolution(None)
    back = batch(calls)
    for p, b in zip(plain, back):
        np.testing.assert_array_equal(p, b)
    batch.release()


def test_smearing_rejects_bad_bands():
    from sasmodels_amd._lib import HipError
    hm = hip_model("sphere")
    q = np.logspace(-3, -1, 40)
    batch = HipBatch(hm, [[q]])
    W = np.zeros((41, 40))               # one row more than the set has q points
    bad = type("Res", (), {"weight_matrix": W})()
    with pytest.raises(ValueError):
        batch.set_resolution([bad])
    # the C ABI checks every band against the set's q count and the weights
    from sasmodels_amd import _lib
    n_out = np.array([1], np.int32)
    start, length = np.array([35], np.int32), np.array([6], np.int32)   # rows 35..40 of 40
    offset, w = np.array([0], np.int64), np.ones(6)
    with pytest.raises(HipError):
        _lib.check(_lib.lib().saship_batch_set_resolution(batch._handle, n_out.ctypes.data, start.ctypes.data,
                                                          length.ctypes.data, None, offset.ctypes.data,
                                                          w.ctypes.data, 6))
    stride = np.array([8], np.int64)                                    # rows 0, 8, .., 40 of 40
    start, length = np.array([0], np.int32), np.array([6], np.int32)
    with pytest.raises(HipError):
        _lib.check(_lib.lib().saship_batch_set_resolution(batch._handle, n_out.ctypes.data, start.ctypes.data,
                                                          length.ctypes.data, stride.ctypes.data,
                                                          offset.ctypes.data, w.ctypes.data, 6))
    # repeated bands: rows 0..4 at repeats 0, 9, 18, 27, 36 reach row 40 of 40
    start, length, stride = np.array([0], np.int32), np.array([5], np.int32), np.array([1], np.int64)
    rep, rstep, roff, rw = np.array([5], np.int32), np.array([9], np.int64), np.array([0], np.int64), np.ones(5)
    args = [n_out, start, length, stride, offset, w]
    with pytest.raises(HipError):
        _lib.check(_lib.lib().saship_batch_set_resolution_repeat(
            batch._handle, *[a.ctypes.data for a in args], 6, rep.ctypes.data, rstep.ctypes.data,
            roff.ctypes.data, rw.ctypes.data, 5))
    rstep[0] = 8                          # ..., 32..36: inside; but repeat weights past the array
    with pytest.raises(HipError):
        _lib.check(_lib.lib().saship_batch_set_resolution_repeat(
            batch._handle, *[a.ctypes.data for a in args], 6, rep.ctypes.data, rstep.ctypes.data,
            roff.ctypes.data, rw.ctypes.data, 4))
    _lib.check(_lib.lib().saship_batch_set_resolution_repeat(
        batch._handle, *[a.ctypes.data for a in args], 6, rep.ctypes.data, rstep.ctypes.data,
        roff.ctypes.data, rw.ctypes.data, 5))
    got = batch(make_calls(hm, q))[0][0]      # the wrapper still splits by q count: the first value
    batch.set_resolution(None)
    plain = batch(make_calls(hm, q))[0]
    np.testing.assert_allclose(got, sum(plain[8*j:8*j + 5].sum() for j in range(5)), rtol=1e-13)
    W = np.eye(40)
    W[39, 0] = 0.5                        # band of point 0 spans all rows: still inside
    ok = type("Res", (), {"weight_matrix": W})()
    batch.set_resolution([ok])
    out = batch(make_calls(hm, q))[0]
    batch.set_resolution(None)
    plain = batch(make_calls(hm, q))[0]
    np.testing.assert_allclose(out, plain @ W, rtol=1e-12)
    batch.release()


def make_calls(hm, q):
    kl = type("K", (), {"info": hm.info, "dtype": np.dtype("d")})()
    return [make_kernel_args(kl, direct_model.get_mesh(hm.info, dict(radius=30), dim="1d"))[:2]]


class _Det:
    def __init__(self, n, span, shift):
        x = np.linspace(-span, span, n)
        qx, qy = np.meshgrid(x + shift, x - 0.5*shift)
        self.qx_data, self.qy_data = qx.ravel(), qy.ravel()
        self.q_data = np.hypot(self.qx_data, self.qy_data)
        self.dqx_data = 0.06*self.q_data + 0.002
        self.dqy_data = 0.03*self.q_data + 0.001


def test_batch_pinhole2d():
    hm = hip_model("cylinder")
    rng = np.random.default_rng(5)
    kl = type("K", (), {"info": hm.info, "dtype": np.dtype("d")})()
    res, calls = [], []
    for s, acc in enumerate(("low", "med", "low")):
        res.append(R.Pinhole2D(_Det(12 + 2*s, 0.2, 0.003*(s + 1)), accuracy=acc))
        pars = dict(radius=rng.uniform(20, 60), length=rng.uniform(100, 400), theta=rng.uniform(0, 90),
                    phi=rng.uniform(0, 90), theta_pd=5, theta_pd_n=3)
        calls.append(make_kernel_args(kl, direct_model.get_mesh(hm.info, pars, dim="2d"))[:2])
    batch = HipBatch(hm, [list(r.q_calc) for r in res])
    plain = batch(calls)
    batch.set_resolution(res)
    smeared = batch(calls)
    for r, p, sm, (det, val) in zip(res, plain, smeared, calls):
        print("pinhole2d", len(r.qx_data), flush=True)
        np.testing.assert_allclose(sm, r.apply(p), rtol=1e-12, atol=0)
        ref = r.apply(OracleKernel(hm.info, list(r.q_calc))(det, val, 0.0, False))
        assert masked_relerr(sm, ref) <= RTOL
    batch.release()


def test_batch_slit2d():
    """Oriented samples measured along qx (Slit2D): the (qy, qx) grid through
    the 2D kernel, the qx bands repeated over the qy rows on the device."""
    from sasmodels_amd import resolution2d
    hm = hip_model("cylinder")
    rng = np.random.default_rng(8)
    kl = type("K", (), {"info": hm.info, "dtype": np.dtype("d")})()
    res, calls = [], []
    for s, (length, acc) in enumerate(((0.0, "low"), (0.004, "low"), (0.002, "med"))):
        q = np.linspace(0.004, 0.12, 24 + 8*s)
        kw = dict(q_calc=q) if length == 0.0 else {}
        res.append(resolution2d.Slit2D(q, q_length=length, q_width=0.01 + 0.005*s, accuracy=acc, **kw))
        pars = dict(radius=rng.uniform(20, 60), length=rng.uniform(100, 400), theta=rng.uniform(0, 90),
                    phi=rng.uniform(0, 90), theta_pd=5, theta_pd_n=3)
        calls.append(make_kernel_args(kl, direct_model.get_mesh(hm.info, pars, dim="2d"))[:2])
    batch = HipBatch(hm, [list(r.q_calc) for r in res])
    plain = batch(calls)
    batch.set_resolution(res)
    smeared = batch(calls)
    for r, p, sm, (det, val) in zip(res, plain, smeared, calls):
        print("slit2d", r.nx, r.ny, flush=True)
        assert len(sm) == len(r.q)
        np.testing.assert_allclose(sm, r.apply(p), rtol=1e-12, atol=0)
        ref = r.apply(OracleKernel(hm.info, list(r.q_calc))(det, val, 0.0, False))
        assert masked_relerr(sm, ref) <= RTOL
    # mixed with a set that has no repeated bands
    batch.set_resolution([res[0], None, res[2]])
    mixed = batch(calls)
    np.testing.assert_allclose(mixed[0], smeared[0], rtol=1e-14)
    np.testing.assert_array_equal(mixed[1], plain[1])
    batch.release()
