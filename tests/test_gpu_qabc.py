"""
The (theta, phi, psi) shapes -- parallelepiped, triaxial ellipsoid and the
bcc, fcc and sc paracrystals -- on
the batched persistent path: nested-rule 1D sets (adaptive and fixed
Gauss-76) and oriented, jittered, magnetic 2D sets, each against the FP64
oracle on the same buffers.  (The single-call path is covered with the
other models in test_gpu_parity.py.)
"""

import numpy as np
import pytest

from gpu_helpers import RTOL, compare_raw, hip_model, masked_relerr
from oracle.oracle import OracleKernel
from sasmodels_amd import direct_model
from sasmodels_amd.details import make_kernel_args
from sasmodels_amd.kernelhip import HipBatch

pytestmark = pytest.mark.gpu


def _sets_1d(model, n, rng):
    out = []
    for _ in range(n):
        q = np.logspace(-3, 0, int(rng.integers(20, 90)))
        if model == "parallelepiped":
            pars = dict(length_a=rng.uniform(10, 80), length_b=rng.uniform(20, 150), length_c=rng.uniform(50, 500),
                        length_b_pd=0.1, length_b_pd_n=int(rng.integers(1, 5)))
        else:
            pars = dict(radius_equat_minor=rng.uniform(10, 100), radius_equat_major=rng.uniform(20, 200),
                        radius_polar=rng.uniform(10, 150), radius_equat_major_pd=0.1,
                        radius_equat_major_pd_n=int(rng.integers(1, 5)))
        out.append((q, pars))
    return out


@pytest.mark.parametrize("model", ["parallelepiped", "triaxial_ellipsoid"])
@pytest.mark.parametrize("quad", [0, 76])
def test_batch_1d(model, quad):
    rng = np.random.default_rng(40 + quad + len(model))
    hm = hip_model(model, quad)
    sets = _sets_1d(model, 6, rng)
    kl = type("K", (), {"info": hm.info, "dtype": np.dtype("d")})()
    calls = [make_kernel_args(kl, direct_model.get_mesh(hm.info, p, dim="1d"))[:2] for _, p in sets]
    batch = HipBatch(hm, [[q] for q, _ in sets])
    print("qabc batch 1d", model, quad, flush=True)
    raw, iq = batch.raw(calls, 0.0, False, 1)
    off = 0
    for s, ((q, _), (det, val)) in enumerate(zip(sets, calls)):
        r = raw[off:off + 2*q.size + 4]
        off += 2*q.size + 4
        ok = OracleKernel(hm.info, [q])
        compare_raw(r, ok.raw(det, val, 0.0, False, 1), q.size, 2)
        assert masked_relerr(iq[batch.qoff[s]:batch.qoff[s + 1]], ok(det, val, 0.0, False)) <= RTOL
    batch.release()


@pytest.mark.parametrize("model", ["parallelepiped", "triaxial_ellipsoid"])
def test_batch_2d_magnetic(model):
    rng = np.random.default_rng(77 + len(model))
    hm = hip_model(model)
    x = np.linspace(-0.2, 0.2, 19)
    qx, qy = [a.ravel() for a in np.meshgrid(x, x + 0.002)]
    kl = type("K", (), {"info": hm.info, "dtype": np.dtype("d")})()
    sizes = (("length_a", "length_b", "length_c") if model == "parallelepiped"
             else ("radius_equat_minor", "radius_equat_major", "radius_polar"))
    calls = []
    for _ in range(4):
        pars = {k: rng.uniform(15, 150) for k in sizes}
        pars.update(theta=rng.uniform(0, 90), phi=rng.uniform(0, 180), psi=rng.uniform(0, 180),
                    psi_pd=10, psi_pd_n=3, phi_pd=5, phi_pd_n=3, sld_M0=rng.uniform(0, 3),
                    sld_mtheta=rng.uniform(0, 90), sld_mphi=rng.uniform(0, 360), up_frac_i=rng.uniform(0, 1),
                    up_frac_f=rng.uniform(0, 1), up_theta=rng.uniform(0, 90))
        det, val, mag = make_kernel_args(kl, direct_model.get_mesh(hm.info, pars, dim="2d"))
        assert mag
        calls.append((det, val))
    batch = HipBatch(hm, [[qx, qy]]*len(calls))
    print("qabc batch 2d magnetic", model, flush=True)
    raw, _ = batch.raw(calls, 0.0, True, 0)
    nq = qx.size
    for s, (det, val) in enumerate(calls):
        ora = OracleKernel(hm.info, [qx, qy]).raw(det, val, 0.0, True, 0)
        compare_raw(raw[s*(nq + 4):(s + 1)*(nq + 4)], ora, nq, 1)
    batch.release()


@pytest.mark.parametrize("model", ["bcc_paracrystal", "fcc_paracrystal", "sc_paracrystal"])
def test_batch_paracrystal_1d(model):
    """FP64 product-rule sets through the persistent batch: <F^2> only, one
    output per q in the raw result."""
    rng = np.random.default_rng(90 + len(model))
    hm = hip_model(model)
    kl = type("K", (), {"info": hm.info, "dtype": np.dtype("d")})()
    sets = []
    for _ in range(4):
        q = np.logspace(-3, -0.5, int(rng.integers(10, 40)))
        pars = dict(dnn=rng.uniform(100, 300), d_factor=rng.uniform(0.03, 0.3), radius=rng.uniform(10, 50),
                    radius_pd=0.1, radius_pd_n=int(rng.integers(1, 4)))
        sets.append((q, make_kernel_args(kl, direct_model.get_mesh(hm.info, pars, dim="1d"))[:2]))
    batch = HipBatch(hm, [[q] for q, _ in sets])
    print("paracrystal batch 1d", model, flush=True)
    raw, iq = batch.raw([c for _, c in sets], 0.0, False, 0)
    off = 0
    for s, (q, (det, val)) in enumerate(sets):
        r = raw[off:off + q.size + 4]
        off += q.size + 4
        ok = OracleKernel(hm.info, [q])
        compare_raw(r, ok.raw(det, val, 0.0, False, 0), q.size, 1)
        assert masked_relerr(iq[batch.qoff[s]:batch.qoff[s + 1]], ok(det, val, 0.0, False)) <= RTOL
    batch.release()
