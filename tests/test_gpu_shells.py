"""
GPU parity of the concentric-shell models with vector parameters
(core_multi_shell: FP32; onion: FP64 like the reference), against the FP64
oracle on identical buffers: 1D with every shell count 0..10 and
polydisperse radius and thicknesses, 2D, magnetic (an M0 on core, solvent
and shell slots), the batched path over sets of different shell counts,
and the Fq outputs with both effective-radius modes.
"""

import zlib

import numpy as np
import pytest

from gpu_helpers import RTOL, buffers, compare_raw, hip_model, masked_relerr
from oracle.oracle import OracleKernel
from sasmodels_amd import direct_model
from sasmodels_amd.details import make_kernel_args
from sasmodels_amd.kernelhip import HipBatch

pytestmark = pytest.mark.gpu

MODELS = ("core_multi_shell", "onion")


def _shells(model, r, n, pd=True):
    """Random parameters with n shells (n may be fractional: rounded by the model)."""
    ctl, core_r = ("n", "radius") if model == "core_multi_shell" else ("n_shells", "radius_core")
    pars = {ctl: n, core_r: r.uniform(10, 120), "sld_core": r.uniform(-1, 7), "sld_solvent": r.uniform(-1, 7)}
    for k in range(1, 11):
        pars["thickness%d" % k] = r.uniform(2, 40)
        if model == "core_multi_shell":
            pars["sld%d" % k] = r.uniform(-1, 7)
        else:
            pars["sld_in%d" % k] = r.uniform(-1, 7)
            pars["sld_out%d" % k] = r.uniform(-1, 7)
            pars["A%d" % k] = float(r.choice([0.0, 1e-4, -2e-3, 0.7, -1.5, 3.0, r.uniform(-4, 4)]))
    if pd:
        pars.update({core_r + "_pd": 0.1, core_r + "_pd_n": int(r.integers(1, 8)),
                     "thickness2_pd": 0.15, "thickness2_pd_n": 3, "thickness1_pd": 0.1, "thickness1_pd_n": 2})
    return pars


@pytest.mark.parametrize("model", MODELS)
def test_shells_1d_all_counts(model):
    r = np.random.default_rng(300 + zlib.crc32(model.encode()) % 1000)
    q = np.concatenate([[0.0], np.logspace(-3, 0, 96)])
    for n in list(range(11)) + [2.3, 3.6]:
        pars = _shells(model, r, n)
        print("1d", model, n, flush=True)
        hk, ok, det, val, mag = buffers(model, [q], pars)
        hk._call_kernel(det, val, 0.0, mag, 1)
        compare_raw(hk.result, ok.raw(det, val, 0.0, mag, 1), q.size, 2)
        hk.release()


@pytest.mark.parametrize("model", MODELS)
def test_shells_2d(model):
    r = np.random.default_rng(400 + zlib.crc32(model.encode()) % 1000)
    x = np.linspace(-0.3, 0.3, 24)
    qv = [a.ravel() for a in np.meshgrid(x, x)]
    for n in (0, 1, 4, 10):
        pars = _shells(model, r, n)
        print("2d", model, n, flush=True)
        hk, ok, det, val, mag = buffers(model, qv, pars)
        hk._call_kernel(det, val, 0.0, mag, 0)
        compare_raw(hk.result, ok.raw(det, val, 0.0, mag, 0), qv[0].size, 1)
        hk.release()


@pytest.mark.parametrize("model", MODELS)
def test_shells_magnetic(model):
    r = np.random.default_rng(500 + zlib.crc32(model.encode()) % 1000)
    x = np.linspace(-0.3, 0.3, 20)
    qv = [a.ravel() for a in np.meshgrid(x, x)]
    qv[0][0] = qv[1][0] = 0.0
    shell = "sld3" if model == "core_multi_shell" else "sld_out3"
    for n in (3, 6):
        pars = _shells(model, r, n, pd=False)
        pars.update({"sld_core_M0": 2.0, "sld_core_mtheta": 30, "sld_solvent_M0": 0.8, "sld_solvent_mphi": 40,
                     shell + "_M0": 3.0, shell + "_mtheta": 70, shell + "_mphi": 200,
                     "up_frac_i": 0.3, "up_frac_f": 0.8, "up_theta": 40, "up_phi": 10})
        print("magnetic", model, n, flush=True)
        hk, ok, det, val, mag = buffers(model, qv, pars)
        assert mag
        hk._call_kernel(det, val, 0.0, mag, 0)
        compare_raw(hk.result, ok.raw(det, val, 0.0, mag, 0), qv[0].size, 1)
        assert hk.result[0] == 0.0
        hk.release()


@pytest.mark.parametrize("model", MODELS)
def test_shells_batch_mixed_counts(model):
    m = hip_model(model)
    r = np.random.default_rng(600 + zlib.crc32(model.encode()) % 1000)
    sets = []
    for s in range(12):
        q = np.logspace(-3, np.log10(0.5), int(r.integers(40, 300)))
        sets.append(([q], _shells(model, r, s % 11)))
    print("batch", model, flush=True)
    batch = HipBatch(m, [qv for qv, _ in sets])
    calls = []
    for qv, pars in sets:
        k = m.make_kernel(qv)
        det, val, _ = make_kernel_args(k, direct_model.get_mesh(m.info, dict(pars), dim="1d"))
        calls.append((det, val))
        k.release()
    raw, iq = batch.raw(calls, 0.0, False, 1)
    off = 0
    for s, ((qv, _), (det, val)) in enumerate(zip(sets, calls)):
        nq = qv[0].size
        ok = OracleKernel(m.info, qv)
        compare_raw(raw[off:off + 2*nq + 4], ok.raw(det, val, 0.0, False, 1), nq, 2)
        off += 2*nq + 4
        assert masked_relerr(iq[batch.qoff[s]:batch.qoff[s + 1]], ok(det, val, 0.0, False)) <= RTOL
    batch.release()


@pytest.mark.parametrize("model", MODELS)
def test_shells_fq_and_radius_modes(model):
    r = np.random.default_rng(700 + zlib.crc32(model.encode()) % 1000)
    q = np.logspace(-3, -0.5, 50)
    pars = _shells(model, r, 5)
    hk, ok, det, val, mag = buffers(model, [q], pars)
    for mode in (1, 2):
        F1, F2, R, V, Vr = hk.Fq(det, val, 0.0, mag, mode)
        G1, G2, S, W, Wr = ok.Fq(det, val, 0.0, mag, mode)
        assert masked_relerr(F2, G2) <= RTOL
        assert np.max(np.abs(F1 - G1)) <= RTOL*np.sqrt(np.max(np.abs(G2)))
        assert abs(R - S) <= 1e-12*abs(S) and abs(V - W) <= 1e-12*abs(W) and abs(Vr - Wr) <= 1e-12
    hk.release()
