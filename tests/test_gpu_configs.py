"""
The benchmark configurations at full size against the FP64 oracle:

  sphere 1D (200 q, radius pd 0.1 x 35), cylinder 1D (1000 q, 35 x 35
  lognormal, Gauss-76), ellipsoid 2D (256 x 256 pixels, 20 x 20 x 20 mesh,
  compared on a 64 x 64 sub-grid and 256 random pixels), batched
  core-shell cylinders with uneven q counts and meshes, and magnetic
  core-shell spheres on a 128 x 128 detector.

Tolerance: 5e-5 relative on I(q) where I > 1e-6 max I.
"""

import os

import numpy as np
import pytest

from gpu_helpers import RTOL, hip_model, masked_relerr
from oracle import oracle
from oracle.oracle import OracleKernel
from sasmodels_amd import direct_model
from sasmodels_amd.details import make_kernel_args
from sasmodels_amd.kernelhip import HipBatch

pytestmark = pytest.mark.gpu
SEED = 20261203


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    oracle.set_threads(min(16, len(os.sched_getaffinity(0))))


def _run(model, qv, pars, quad=0, magnetic=None):
    hm = hip_model(model, quad)
    k = hm.make_kernel(qv)
    mesh = direct_model.get_mesh(hm.info, dict(pars), dim=k.dim)
    det, val, mag = make_kernel_args(k, mesh)
    mag = mag if magnetic is None else magnetic
    I = k(det, val, 0.0, mag)
    k.release()
    return hm, det, val, mag, I


def test_sphere_1d():
    q = np.logspace(-3, np.log10(0.5), 200)
    pars = dict(radius=50, radius_pd=0.1, radius_pd_n=35, radius_pd_nsigma=3, sld=4, sld_solvent=1,
                scale=1, background=1e-3)
    hm, det, val, mag, I = _run("sphere", [q], pars)
    ref = OracleKernel(hm.info, [q])(det, val, 0.0, mag)
    err = masked_relerr(I, ref)
    print("sphere 1D max rel err %.3g" % err)
    assert err <= RTOL


def test_cylinder_1d_gauss76():
    q = np.logspace(-3, 0, 1000)
    pars = dict(radius=20, length=400, sld=4, sld_solvent=1,
                radius_pd=0.1, radius_pd_n=35, radius_pd_type="lognormal",
                length_pd=0.1, length_pd_n=35, length_pd_type="lognormal")
    hm, det, val, mag, I = _run("cylinder", [q], pars, quad=76)
    assert det.num_eval == 1225
    ref = OracleKernel(hm.info, [q])(det, val, 0.0, mag)
    err = masked_relerr(I, ref)
    print("cylinder 1D max rel err %.3g" % err)
    assert err <= RTOL


def test_ellipsoid_2d_256():
    x = np.linspace(-0.3, 0.3, 256)
    qx, qy = [a.ravel() for a in np.meshgrid(x, x)]
    pars = dict(radius_polar=20, radius_equatorial=400, theta=60, phi=60,
                theta_pd=10, theta_pd_n=20, phi_pd=20, phi_pd_n=20,
                radius_equatorial_pd=0.1, radius_equatorial_pd_n=20)
    hm, det, val, mag, I = _run("ellipsoid", [qx, qy], pars)
    assert det.num_eval == 8000
    sub = np.arange(0, 256, 4)
    idx = (sub[:, None]*256 + sub[None, :]).ravel()
    rng = np.random.default_rng(SEED)
    idx = np.unique(np.concatenate([idx, rng.integers(0, 256*256, 256)]))
    ref = OracleKernel(hm.info, [qx[idx], qy[idx]])(det, val, 0.0, mag)
    err = masked_relerr(I[idx], ref)
    print("ellipsoid 2D max rel err %.3g on %d pixels" % (err, idx.size))
    assert err <= RTOL


def test_batched_core_shell_cylinder():
    rng = np.random.default_rng(SEED)
    hm = hip_model("core_shell_cylinder", 76)
    n = 64
    q_list, calls = [], []
    kl = type("K", (), {"info": hm.info, "dtype": np.dtype("d")})()
    for _ in range(n):
        nq = int(rng.integers(64, 2049))
        q = np.logspace(-3, np.log10(0.5), nq)
        pars = dict(radius=rng.uniform(10, 100), thickness=rng.uniform(5, 30), length=rng.uniform(50, 1000),
                    sld_core=rng.uniform(-1, 7), sld_shell=rng.uniform(-1, 7), sld_solvent=rng.uniform(-1, 7),
                    radius_pd=rng.uniform(0, 0.2), radius_pd_n=int(rng.integers(1, 36)))
        mesh = direct_model.get_mesh(hm.info, pars, dim="1d")
        q_list.append([q])
        calls.append(make_kernel_args(kl, mesh)[:2])
    batch = HipBatch(hm, q_list)
    out = batch(calls)
    worst = 0.0
    for qv, (det, val), I in zip(q_list, calls, out):
        ref = OracleKernel(hm.info, qv)(det, val, 0.0, False)
        worst = max(worst, masked_relerr(I, ref))
    batch.release()
    print("batched core_shell_cylinder worst max rel err %.3g" % worst)
    assert worst <= RTOL


def test_magnetic_core_shell_sphere_128():
    rng = np.random.default_rng(SEED + 1)
    hm = hip_model("core_shell_sphere")
    x = np.linspace(-0.3, 0.3, 128)
    qx, qy = [a.ravel() for a in np.meshgrid(x, x)]
    kl = type("K", (), {"info": hm.info, "dtype": np.dtype("d")})()
    calls = []
    for _ in range(4):
        pars = dict(sld_core_M0=rng.uniform(0, 3), sld_shell_M0=rng.uniform(0, 3),
                    sld_solvent_M0=rng.uniform(0, 3),
                    sld_core_mtheta=rng.uniform(0, 90), sld_core_mphi=rng.uniform(0, 360),
                    sld_shell_mtheta=rng.uniform(0, 90), sld_shell_mphi=rng.uniform(0, 360),
                    sld_solvent_mtheta=rng.uniform(0, 90), sld_solvent_mphi=rng.uniform(0, 360),
                    up_frac_i=rng.uniform(0, 1), up_frac_f=rng.uniform(0, 1), up_theta=rng.uniform(0, 90))
        mesh = direct_model.get_mesh(hm.info, pars, dim="2d")
        det, val, mag = make_kernel_args(kl, mesh)
        assert mag
        calls.append((det, val))
    batch = HipBatch(hm, [[qx, qy]]*len(calls))
    out = batch(calls, magnetic=True)
    worst = 0.0
    for (det, val), I in zip(calls, out):
        ref = OracleKernel(hm.info, [qx, qy])(det, val, 0.0, True)
        worst = max(worst, masked_relerr(I, ref))
    batch.release()
    print("magnetic core_shell_sphere 128x128 worst max rel err %.3g" % worst)
    assert worst <= RTOL


def test_core_shell_cylinder_full_4096_launch():
    """The whole 4096-set launch of the benchmark (uneven q counts and meshes,
    one persistent launch): every set finite and positive; launch-shape
    invariance over all sets (another item size and grid only reorders FP32
    sums); the single-call path on 64 sets and the FP64 oracle on 16 sets
    spread over the launch; the forward intensity of every set against its
    Guinier estimate from the oracle's own first q."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    from sasmodels_amd import core
    wl = bench.wl_core_shell_cylinder(4096)
    model = core.build_model(wl["info"], platform="hip")
    batch = HipBatch(model, wl["q_list"])
    calls = wl["calls"]
    batch.set_options(pts_per_item=16, waves_per_cu=32)
    first = batch(calls)
    batch.set_options(pts_per_item=4, waves_per_cu=16)
    second = batch(calls)
    batch.release()
    assert len(first) == 4096
    for a, b in zip(first, second):
        assert np.all(np.isfinite(a)) and np.all(a > 0)
        assert masked_relerr(b, a) <= 2e-6
    rng = np.random.default_rng(SEED + 9)
    pick = rng.choice(4096, 64, replace=False)
    for i in pick:
        k = model.make_kernel(wl["q_list"][i])
        single = k(*calls[i], 0.0, False)
        k.release()
        assert masked_relerr(first[i], single) <= 2e-6
    worst = 0.0
    for i in pick[:16]:
        ref = OracleKernel(wl["info"], wl["q_list"][i])(*calls[i], 0.0, False)
        worst = max(worst, masked_relerr(first[i], ref))
    model.release()
    print("4096-set launch: worst rel err vs FP64 over 16 sets %.3g" % worst)
    assert worst <= RTOL
