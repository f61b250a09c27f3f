"""
Benchmark of the I(q) path on one MI355X.

Headline workload (the cylinder configuration the metric is quoted on):
cylinder 1D, 1000 log-spaced q in [1e-3, 1] 1/A, radius 20 A and length
400 A each with lognormal polydispersity 0.1 (35 x 35 mesh points, 3
sigma), sld 4 / solvent 1, orientation average by fixed Gauss-76, FP32
kernels with compensated sums and FP64 weight sums, cutoff 0.

Paths, one JSON line:
  batched          (value)  a step evaluates BATCH curves of the workload in
                            one persistent launch (separable cylinder path)
                            with its inputs resident in HBM (device work
                            only, replayed back to back); the host-inclusive
                            rate -- a full call per step with new jittered
                            inputs: packing, upload, graph replay, download
                            -- is reported beside it
  batched_cutoff            the same at cutoff 1e-5 (pruned mesh corners)
  batched_generic           the same on the generic mesh kernel
  plain, plain_generic      one curve per reference-shaped call
  surrogate                 the trained MLP on 4096 curves x 512 q drawn from
                            its box (fused inference, inputs in HBM), against
                            the plain batched kernel on 256 of those curves
                            (device-resident replays)
roofline: FLOPs of the separable evaluation (counted with the reference's
formulas for J1 and sin, see separable_flops_cylinder) over the device time
of k_batch_sep (HIP events on the batch's stream), against the published
FP32 peak and against the FMA rate measured on the device (saship_probe_fma,
non-packed instructions); traffic and VALU instruction counts of the same
launch from the counter summary in profiles/ when present.  cpu_baseline: the FP64 oracle on the
host cores, on a bounded sample of the workload.

other_configs: the sphere 1D, pinhole-smeared cylinder, cylinder@hardsphere
(beta), ellipsoid 2D, batched core-shell cylinder, batched magnetic
core-shell sphere, batched parallelepiped (nested orientation rules) and
batched bcc paracrystal (FP64) configurations, each with device throughput, FP64 accuracy on a subset and
the oracle's rate on the host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu]
                    [--no-surrogate] [--no-configs]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "I(q) curves/s: plain HIP vs batched vs surrogate, at stated rel err vs FP64"
PEAK_FP32_TFLOPS = 157.3
SEED = 20261203
SURROGATE_WEIGHTS = os.path.join(ROOT, "sasmodels_amd", "data", "cylinder_surrogate.npz")

C2_PARS = dict(radius=20.0, length=400.0, sld=4.0, sld_solvent=1.0, scale=1.0, background=1e-3,
               radius_pd=0.1, radius_pd_n=35, radius_pd_nsigma=3, radius_pd_type="lognormal",
               length_pd=0.1, length_pd_n=35, length_pd_nsigma=3, length_pd_type="lognormal")
C2_Q = np.logspace(-3, 0, 1000)
QUAD = 76


def reference_flops_cylinder_fq(q, radii):
    """
    FLOPs of one cylinder <F>,<F^2> pass as the reference formulas state it
    (cylinder.c:55-88 with sas_J1.c:107-153, polevl.c, kernel_header.c:177)
    and the pd-loop accumulation (kernel_iq.c:797-799), counting add, sub,
    mul, div, sqrt, sin and cos as 1 each, for the branches taken:
      per Gauss node: 21 + J1 (29 for |x| <= 5, 67 beyond)
      per (q, mesh point): 18
    *radii* holds the radius of every mesh point.
    """
    z, _ = np.polynomial.legendre.leggauss(QUAD)
    s = np.sin(z*np.pi/4 + np.pi/4)
    R = np.asarray(radii, "d")
    total = 0.0
    for qi in q:
        x = qi*R[:, None]*s[None, :]
        small = np.count_nonzero(np.abs(x) <= 5.0)
        total += small*(21 + 29) + (x.size - small)*(21 + 67) + 18*R.size
    return total


def separable_flops_cylinder(q, radii, nl):
    """
    FLOPs of the separable evaluation of one curve with the same convention
    (add, sub, mul, div, sin = 1, FMA = 2; J1 with the reference's formula,
    29 for |x| <= 5, 67 beyond), prefix mode with every row complete:
      per (q, node): 6 + 8 per length point + per radius point (10 + J1)
    (q cos/sin products and weighted sums; per length point the sinc and
    two prefix sums; per radius point the J1 argument, 2J1/x and two
    products into F and F^2).  *radii* holds the distinct radii.
    """
    z, _ = np.polynomial.legendre.leggauss(QUAD)
    s = np.sin(z*np.pi/4 + np.pi/4)
    R = np.asarray(radii, "d")
    total = 0.0
    for qi in q:
        x = qi*R[:, None]*s[None, :]
        small = np.count_nonzero(np.abs(x) <= 5.0)
        total += s.size*(6 + 8*nl) + small*(10 + 29) + (x.size - small)*(10 + 67)
    return total


def _kernel_like(info):
    return type("K", (), {"info": info, "dtype": np.dtype("d")})()


def c2_calls(info, n, offset=0, extra=None):
    """(details, values) for n jittered copies of the workload (plus *extra* parameters)."""
    from sasmodels_amd import direct_model
    from sasmodels_amd.details import make_kernel_args
    kl = _kernel_like(info)
    out = []
    for k in range(offset, offset + n):
        f = 1 + 1e-4*(k % 97)
        pars = dict(C2_PARS, radius=C2_PARS["radius"]*f, length=C2_PARS["length"]*(2 - f), **(extra or {}))
        out.append(make_kernel_args(kl, direct_model.get_mesh(info, pars, dim="1d"))[:2])
    return out


def masked_relerr(a, b):
    sel = b > 1e-6*b.max()
    return np.abs(a[sel] - b[sel])/b[sel]


def err_stats(rel):
    rel = np.concatenate([np.ravel(r) for r in rel])
    return {"max": float(rel.max()), "median": float(np.median(rel)),
            "p98": float(np.percentile(rel, 98)), "rms": float(np.sqrt(np.mean(rel**2)))}


def accuracy(out, refs):
    """Relative error of GPU curves against the FP64 oracle curves (pairs in order)."""
    return err_stats([masked_relerr(np.asarray(o), r) for o, r in zip(out, refs)])


def cpu_baseline(info, jobs, threads, budget_s=12.0):
    """
    The CPU leg: the FP64 oracle (C, OpenMP over q) on whole workload curves
    -- *jobs* is a list of (details, values, cutoff) -- in order until
    budget_s of host time is spent, and on one thread on a q subsample.
    Returns the baseline record and the oracle's I(q) of the curves it ran,
    which are the FP64 references for the accuracy figures.
    """
    from oracle import oracle
    from oracle.oracle import OracleKernel
    oracle.set_threads(threads)
    ok = OracleKernel(info, [C2_Q])
    refs, t0 = [], time.perf_counter()
    for det, val, cutoff in jobs:
        if time.perf_counter() - t0 >= budget_s:
            break
        refs.append(ok(det, val, cutoff, False))
    dt = time.perf_counter() - t0
    oracle.set_threads(1)
    det, val, _ = jobs[0]
    sample = C2_Q[::25]
    t1 = time.perf_counter()
    oracle.raw_call("cylinder", 0, QUAD, sample, det.buffer, val)
    dt1 = (time.perf_counter() - t1)*C2_Q.size/sample.size
    oracle.set_threads(threads)
    rec = {"value": len(refs)/dt, "unit": "curves/s", "cores": threads, "kind": "port",
           "sample": "oracle/sasoracle.c FP64, %d whole workload curves (1000 q, 35x35 mesh, Gauss-76) "
                     "in %.1f s on %d OpenMP threads" % (len(refs), dt, threads),
           "single_thread_curves_per_s": 1.0/dt1,
           "single_thread_sample": "1 thread, 40 of the 1000 q of one curve, scaled to a whole curve"}
    return rec, refs


def bench_plain(model, info, steps, warmup, separable=True):
    kernel = model.make_kernel([C2_Q])
    kernel.set_separable(separable)
    calls = c2_calls(info, steps + warmup, offset=1000)
    for det, val in calls[:warmup]:
        kernel(det, val, 0.0, False)
    kernel.set_timing(True)
    t0 = time.perf_counter()
    for det, val in calls[warmup:]:
        kernel(det, val, 0.0, False)
    dt = time.perf_counter() - t0
    ms, n = kernel.get_timing()
    kernel.release()
    return {"curves_per_s": steps/dt, "ms_per_curve": dt/steps*1e3, "eval_kernel_ms": ms/max(n, 1)}


def bench_batched(model, info, steps, warmup, batch_size, separable=True, cutoff=0.0):
    """
    Two timings of a batch of the workload: (1) device-resident, the value:
    `steps` back-to-back runs of the device work (expand, evaluate, final
    sums) with inputs and outputs in HBM, wall clock bracketed by
    synchronisation; (2) host-inclusive: `steps` full calls with different
    jittered inputs each (host packing, staging upload, hipGraph replay,
    result download).
    """
    from sasmodels_amd.kernelhip import HipBatch
    batch = HipBatch(model, [[C2_Q]]*batch_size, use_graph=True)
    batch.set_separable(separable)
    groups = [c2_calls(info, batch_size, offset=g*batch_size) for g in range(steps + warmup)]
    pts, waves = batch.autotune(groups[0], cutoff)
    for calls in groups[:warmup]:
        batch(calls, cutoff)
    t0 = time.perf_counter()
    for calls in groups[warmup:]:
        out = batch(calls, cutoff)
    dt_host = time.perf_counter() - t0
    # device time of the evaluation kernel (eager launches with events)
    batch.set_timing(True)
    for calls in groups[warmup:warmup + min(steps, 10)]:
        batch(calls, cutoff)
    ms, n = batch.get_timing()
    batch.set_timing(False)
    first = batch(groups[0], cutoff)
    batch.replay(max(1, warmup))
    t0 = time.perf_counter()
    dev_ms = batch.replay(steps)
    dt = time.perf_counter() - t0
    batch.release()
    return ({"curves_per_s": steps*batch_size/dt, "ms_per_step": dt/steps*1e3,
             "device_ms_per_step": dev_ms/steps, "batch": batch_size,
             "curves_per_s_host_inclusive": steps*batch_size/dt_host,
             "ms_per_step_host_inclusive": dt_host/steps*1e3,
             "cutoff": cutoff, "separable": separable,
             "pts_per_item": pts, "waves_per_cu": waves, "graph": True,
             "eval_kernel_ms": ms/max(n, 1)}, groups[0], first)


def bench_surrogate(steps=20):
    """Surrogate inference on 4096 curves x 512 q against the plain kernel."""
    if not os.path.exists(SURROGATE_WEIGHTS):
        return None
    import torch
    from sasmodels_amd.surrogate import CylinderSurrogate, sample_box
    sur = CylinderSurrogate()
    sur.load(SURROGATE_WEIGHTS)
    rng = np.random.default_rng(SEED)
    ncurves, nq = 4096, 512
    curves = sample_box(ncurves, rng)
    lq = np.linspace(-3, 0, nq)
    c_t = torch.as_tensor(curves.astype(np.float32), device="cuda")
    q_t = torch.as_tensor(lq.astype(np.float32), device="cuda")
    out = torch.empty(ncurves*nq, dtype=torch.float32, device="cuda")
    sur.repack()
    for _ in range(3):
        sur.infer_into(c_t, q_t, out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        sur.infer_into(c_t, q_t, out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)/steps
    # accuracy and plain-kernel cost on a subset of the same curves
    nsub = 256
    pred = out.view(ncurves, nq)[:nsub].double().cpu().numpy()
    q = 10.0**lq
    b = sur.plain.batch([q]*nsub)
    calls = sur.plain.calls(curves[:nsub])
    b.autotune(calls)
    ref = np.array(b(calls))
    t0 = time.perf_counter()
    b(calls)
    t_host = time.perf_counter() - t0
    b.replay(2)
    t_plain = b.replay(10)*1e-3/10          # device-resident, like the surrogate timing
    b.release()
    err = np.array([masked_relerr(p, r).max() for p, r in zip(pred, ref)])
    flops = 2.0*198400*ncurves*nq
    return {"curves_per_s": ncurves/(ms*1e-3), "ms_per_pass": ms, "curves": ncurves, "nq": nq,
            "tflops_bf16": flops/(ms*1e-3)/1e12,
            "max_rel_err_vs_plain": float(err.max()), "median_rel_err_vs_plain": float(np.median(err)),
            "frac_curves_within_1e-3": float(np.mean(err <= 1e-3)),
            "frac_curves_within_1e-2": float(np.mean(err <= 1e-2)),
            "plain_curves_per_s_same_curves": nsub/t_plain,
            "plain_curves_per_s_same_curves_host_inclusive": nsub/t_host,
            "speedup_vs_plain": (ncurves/(ms*1e-3))/(nsub/t_plain)}


def _timed(fn, reps):
    fn()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    return (time.perf_counter() - t0)/reps, out


def _oracle_leg(info, jobs, threads):
    """CPU leg of a secondary configuration: the oracle on `jobs` =
    [(q_vectors, details, values, magnetic)], timed; returns (seconds, I list)."""
    from oracle import oracle
    from oracle.oracle import OracleKernel
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    refs = [OracleKernel(info, qv)(det, val, 0.0, mag) for qv, det, val, mag in jobs]
    return time.perf_counter() - t0, refs


def bench_configs(threads, cpu=True):
    """
    The other benchmark configurations, one entry each: device throughput
    (batches: device-resident replays, with the host-inclusive call rate
    beside it), the FP64 accuracy and the oracle's time on a bounded subset
    (CPU leg, scaled to the same unit).
    """
    from sasmodels_amd import core, direct_model, generate
    from sasmodels_amd.details import make_kernel_args
    from sasmodels_amd.kernelhip import HipBatch
    rng = np.random.default_rng(SEED)
    out = {}

    # C1: sphere 1D, 200 q, gaussian radius pd 0.1 x 35; 4096 jittered curves per launch
    info = core.load_model_info("sphere")
    model = core.build_model(info, platform="hip")
    q = np.logspace(-3, np.log10(0.5), 200)
    kl = _kernel_like(info)
    calls = [make_kernel_args(kl, direct_model.get_mesh(info, dict(
        radius=50*(1 + 1e-3*(k % 101)), radius_pd=0.1, radius_pd_n=35, radius_pd_nsigma=3, sld=4, sld_solvent=1,
        scale=1, background=1e-3), dim="1d"))[:2] for k in range(4096)]
    b = HipBatch(model, [[q]]*len(calls), use_graph=True)
    b.autotune(calls)
    dt, iq = _timed(lambda: b(calls), 10)
    b.replay(2)
    res = b.replay(10)*1e-3/10
    b.release()
    ent = {"workload": "sphere 1D, 200 q, radius 50 gaussian pd 0.1 (35 points), 4096 curves per launch",
           "curves_per_s": len(calls)/res, "curves_per_s_host_inclusive": len(calls)/dt,
           "ms_per_launch": res*1e3, "qpd_evals_per_s": len(calls)*200*35/res}
    if cpu:
        t, refs = _oracle_leg(info, [([q], d, v, False) for d, v in calls[:64]], threads)
        ent["max_rel_err_vs_fp64"] = float(max(masked_relerr(np.asarray(a), r).max() for a, r in zip(iq, refs)))
        ent["cpu_curves_per_s"] = 64/t
    out["sphere_1d_batched"] = ent
    model.release()

    # cylinder headline curves measured through a pinhole (dq/q = 5 %): the
    # batch evaluates q_calc and smears on the device
    from sasmodels_amd import resolution
    info = core.load_model_info("cylinder")
    generate.set_integration_size(info, QUAD)
    model = core.build_model(info, platform="hip")
    res = resolution.Pinhole1D(C2_Q, 0.05*C2_Q)
    calls = c2_calls(info, 64, offset=9000)
    b = HipBatch(model, [[res.q_calc]]*len(calls), use_graph=True)
    b.autotune(calls)
    b.set_resolution([res]*len(calls))
    dt, iq = _timed(lambda: b(calls), 10)
    b.replay(2)
    with_smear = b.replay(20)*1e-3/20
    b.set_resolution(None)
    b(calls)
    b.replay(2)
    without = b.replay(20)*1e-3/20
    b.release()
    _, length, _, _ = resolution.band(res.weight_matrix)
    smear_bytes = 64*(16*int(length.sum()) + 8*C2_Q.size)
    ent = {"workload": "cylinder headline curves (64 per launch) with pinhole resolution dq/q = 0.05: "
                       "%d q_calc points smeared to 1000 q on the device" % res.q_calc.size,
           "curves_per_s": 64/with_smear, "curves_per_s_host_inclusive": 64/dt,
           "ms_per_launch": with_smear*1e3, "smear_ms": (with_smear - without)*1e3,
           "smear_gbytes_per_s": smear_bytes/max(with_smear - without, 1e-9)/1e9}
    if cpu:
        t, refs = _oracle_leg(info, [([res.q_calc], d, v, False) for d, v in calls[:4]], threads)
        ent["max_rel_err_vs_fp64"] = float(max(masked_relerr(np.asarray(a), res.apply(r)).max()
                                                for a, r in zip(iq, refs)))
    out["cylinder_batched_pinhole"] = ent
    model.release()

    # the cylinder headline curves as cylinder@hardsphere, beta approximation,
    # excluded-volume effective radius: S(q) and the combination fused into
    # the batch's final kernel (FP64)
    from sasmodels_amd.product import ProductBatch
    info = core.load_model_info("cylinder@hardsphere")
    p_info = info.composition[1][0]
    generate.set_integration_size(p_info, QUAD)
    model = core.build_model(info, platform="hip")
    sfac = dict(volfraction=0.2, structure_factor_mode=1, radius_effective_mode=1)
    calls = c2_calls(info, 64, offset=12000, extra=sfac)
    pb = ProductBatch(model, [[C2_Q]]*len(calls), use_graph=True)
    pb.autotune(calls)
    dt, iq = _timed(lambda: pb(calls), 10)
    pb.replay(2)
    with_s = pb.replay(20)*1e-3/20
    p_calls = pb._split(calls)[0]
    pb.batch.set_structure(None)
    pb.batch(p_calls, radius_effective_mode=1)
    pb.replay(2)
    without = pb.replay(20)*1e-3/20
    pb.release()
    ent = {"workload": "cylinder headline curves (64 per launch) times the hard-sphere S(q), volfraction 0.2, "
                       "beta approximation, excluded-volume effective radius",
           "curves_per_s": 64/with_s, "curves_per_s_host_inclusive": 64/dt, "ms_per_launch": with_s*1e3,
           "structure_factor_ms": (with_s - without)*1e3}
    if cpu:
        from oracle import oracle
        from oracle.oracle import OracleKernel
        oracle.set_threads(threads)
        t0 = time.perf_counter()
        refs = []
        for det, val in c2_calls(p_info, 4, offset=12000):
            fq = OracleKernel(p_info, [C2_Q]).Fq(det, val, 0.0, False, 1)
            refs.append(oracle.product(C2_Q, fq, 1.0, 1e-3, 50.0, 0.2, True, 1)[0])
        t = time.perf_counter() - t0
        ent["max_rel_err_vs_fp64"] = float(max(masked_relerr(np.asarray(a), r).max() for a, r in zip(iq, refs)))
        ent["cpu_curves_per_s"] = 4/t
    out["cylinder_hardsphere_beta_batched"] = ent
    model.release()

    # C3: ellipsoid 2D on 256 x 256, theta/phi/radius_equatorial dispersion 20 x 20 x 20
    info = core.load_model_info("ellipsoid")
    model = core.build_model(info, platform="hip")
    x = np.linspace(-0.3, 0.3, 256)
    qx, qy = [a.ravel() for a in np.meshgrid(x, x)]
    pars = dict(radius_polar=20, radius_equatorial=400, theta=60, phi=60, theta_pd=10, theta_pd_n=20,
                phi_pd=20, phi_pd_n=20, radius_equatorial_pd=0.1, radius_equatorial_pd_n=20)
    k = model.make_kernel([qx, qy])
    mesh = direct_model.get_mesh(info, pars, dim="2d")
    det, val, mag = make_kernel_args(k, mesh)
    k.set_timing(True)
    dt, img = _timed(lambda: k(det, val, 0.0, mag), 5)
    ms, n = k.get_timing()
    k.release()
    ent = {"workload": "ellipsoid 2D, 256 x 256 detector, 20 x 20 x 20 mesh (theta, phi, radius_equatorial)",
           "images_per_s": 1.0/dt, "eval_kernel_ms": ms/max(n, 1),
           "qpd_evals_per_s": qx.size*int(det.num_eval)/(ms/max(n, 1)*1e-3)}
    if cpu:
        sub = rng.choice(qx.size, 1024, replace=False)
        t, refs = _oracle_leg(info, [([qx[sub], qy[sub]], det, val, mag)], threads)
        ent["max_rel_err_vs_fp64"] = float(masked_relerr(np.asarray(img)[sub], refs[0]).max())
        ent["max_rel_err_pixels"] = 1024
        ent["cpu_images_per_s"] = 1024/(t*qx.size)
    out["ellipsoid_2d"] = ent
    model.release()

    # C4a: 4096 core-shell cylinder sets, uneven q counts and radius meshes, Gauss-76
    info = core.load_model_info("core_shell_cylinder")
    generate.set_integration_size(info, QUAD)
    model = core.build_model(info, platform="hip")
    kl = _kernel_like(info)
    q_list, calls = [], []
    for _ in range(4096):
        nq = int(rng.integers(64, 2049))
        q_list.append([np.logspace(-3, np.log10(0.5), nq)])
        p = dict(radius=rng.uniform(10, 100), thickness=rng.uniform(5, 30), length=rng.uniform(50, 1000),
                 sld_core=rng.uniform(-1, 7), sld_shell=rng.uniform(-1, 7), sld_solvent=rng.uniform(-1, 7),
                 radius_pd=rng.uniform(0, 0.2), radius_pd_n=int(rng.integers(1, 36)))
        calls.append(make_kernel_args(kl, direct_model.get_mesh(info, p, dim="1d"))[:2])
    b = HipBatch(model, q_list, use_graph=True)
    b.autotune(calls)
    dt, iq = _timed(lambda: b(calls), 3)
    b.set_timing(True)
    b(calls)
    ms, n = b.get_timing()
    b.set_timing(False)
    b(calls)
    b.replay(1)
    res = b.replay(3)*1e-3/3
    b.release()
    evals = sum(qv[0].size*int(d.num_eval) for qv, (d, _) in zip(q_list, calls))
    ent = {"workload": "4096 core_shell_cylinder sets, nq U{64..2048}, radius pd U[0,0.2] x U{1..35} points, "
                       "Gauss-76, one persistent launch", "sets_per_s": 4096/res, "sets_per_s_host_inclusive": 4096/dt,
           "ms_per_launch": res*1e3, "eval_kernel_ms": ms/max(n, 1), "qpd_evals_per_s": evals/(ms/max(n, 1)*1e-3)}
    if cpu:
        t, refs = _oracle_leg(info, [(qv, d, v, False) for qv, (d, v) in zip(q_list[:16], calls[:16])], threads)
        ent["max_rel_err_vs_fp64"] = float(max(masked_relerr(np.asarray(a), r).max() for a, r in zip(iq, refs)))
        e16 = sum(qv[0].size*int(d.num_eval) for qv, (d, _) in zip(q_list[:16], calls[:16]))
        ent["cpu_sets_per_s"] = 4096/(t*evals/e16)
    out["core_shell_cylinder_batched"] = ent
    model.release()

    # C4b: 256 magnetic core-shell sphere sets on a 128 x 128 detector, four spin states
    info = core.load_model_info("core_shell_sphere")
    model = core.build_model(info, platform="hip")
    kl = _kernel_like(info)
    x = np.linspace(-0.3, 0.3, 128)
    qx, qy = [a.ravel() for a in np.meshgrid(x, x)]
    calls = []
    for _ in range(256):
        p = dict(sld_core_M0=rng.uniform(0, 3), sld_shell_M0=rng.uniform(0, 3), sld_solvent_M0=rng.uniform(0, 3),
                 sld_core_mtheta=rng.uniform(0, 90), sld_core_mphi=rng.uniform(0, 360),
                 sld_shell_mtheta=rng.uniform(0, 90), sld_shell_mphi=rng.uniform(0, 360),
                 sld_solvent_mtheta=rng.uniform(0, 90), sld_solvent_mphi=rng.uniform(0, 360),
                 up_frac_i=rng.uniform(0, 1), up_frac_f=rng.uniform(0, 1), up_theta=rng.uniform(0, 90))
        calls.append(make_kernel_args(kl, direct_model.get_mesh(info, p, dim="2d"))[:2])
    b = HipBatch(model, [[qx, qy]]*len(calls), use_graph=True)
    b.autotune(calls, magnetic=True)
    dt, img = _timed(lambda: b(calls, magnetic=True), 3)
    b.replay(1)
    res = b.replay(3)*1e-3/3
    b.release()
    ent = {"workload": "256 magnetic core_shell_sphere sets, 128 x 128 detector, 4 spin states",
           "sets_per_s": 256/res, "sets_per_s_host_inclusive": 256/dt, "ms_per_launch": res*1e3}
    if cpu:
        t, refs = _oracle_leg(info, [([qx, qy], d, v, True) for d, v in calls[:2]], threads)
        ent["max_rel_err_vs_fp64"] = float(max(masked_relerr(np.asarray(a), r).max() for a, r in zip(img, refs)))
        ent["cpu_sets_per_s"] = 2/t
    out["magnetic_core_shell_sphere_batched"] = ent
    model.release()

    # more models on the same machinery: a nested-rule 1D shape and an FP64 one
    for name, nsets, nq, mk, evals_per_point in (
            ("parallelepiped", 256, 200,
             lambda: dict(length_a=rng.uniform(20, 60), length_b=rng.uniform(40, 120),
                          length_c=rng.uniform(100, 400), length_a_pd=0.1, length_a_pd_n=5,
                          length_b_pd=0.1, length_b_pd_n=5), None),
            ("bcc_paracrystal", 64, 200,
             lambda: dict(dnn=rng.uniform(150, 300), d_factor=rng.uniform(0.05, 0.2), radius=rng.uniform(20, 50),
                          radius_pd=0.1, radius_pd_n=5), 150*150)):
        info = core.load_model_info(name)
        model = core.build_model(info, platform="hip")
        kl = _kernel_like(info)
        q = np.logspace(-3, np.log10(0.5), nq)
        calls = [make_kernel_args(kl, direct_model.get_mesh(info, mk(), dim="1d"))[:2] for _ in range(nsets)]
        b = HipBatch(model, [[q]]*nsets, use_graph=True)
        b.autotune(calls)
        dt, iq = _timed(lambda: b(calls), 3)
        b.replay(1)
        res = b.replay(5)*1e-3/5
        b.release()
        npts = sum(int(d.num_eval) for d, _ in calls)
        ent = {"workload": "%d %s sets, %d q in [1e-3, 0.5], %d mesh points in all, one persistent launch%s"
                           % (nsets, name, nq, npts, "" if evals_per_point is None else
                              "; FP64, Gauss-150 x 150 lattice average per (q, point)"),
               "curves_per_s": nsets/res, "curves_per_s_host_inclusive": nsets/dt, "ms_per_launch": res*1e3,
               "qpd_evals_per_s": nq*npts/res}
        if evals_per_point:
            ent["lattice_evals_per_s"] = nq*npts*evals_per_point/res
        if cpu:
            sub = 4
            t, refs = _oracle_leg(info, [([q], d, v, False) for d, v in calls[:sub]], threads)
            ent["max_rel_err_vs_fp64"] = float(max(masked_relerr(np.asarray(a), r).max()
                                                    for a, r in zip(iq, refs)))
            ent["cpu_curves_per_s"] = sub/t
        out[name + "_batched"] = ent
        model.release()
    if cpu:
        for e in out.values():
            e["cpu_threads"] = threads
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the other benchmark configurations")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a kernel of known traffic for "
                         "scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from sasmodels_amd import core, direct_model, generate

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy*(1 << 20)//4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    threads = min(16, len(os.sched_getaffinity(0)))
    info = core.load_model_info("cylinder")
    generate.set_integration_size(info, QUAD)
    model = core.build_model(info, platform="hip")

    batched, calls0, out0 = bench_batched(model, info, args.steps, args.warmup, args.batch)
    short = max(2, min(args.steps, 10))
    cut, ccalls, cout = bench_batched(model, info, args.steps, args.warmup, args.batch, cutoff=1e-5)
    generic, gcalls, gout = bench_batched(model, info, short, 1, args.batch, separable=False)
    plain = bench_plain(model, info, args.steps, args.warmup)
    plain_generic = bench_plain(model, info, short, 1, separable=False)

    mesh = direct_model.get_mesh(info, C2_PARS, dim="1d")
    radii, lengths = mesh[4][1], mesh[5][1]
    npts = radii.size*lengths.size
    flops_curve = separable_flops_cylinder(C2_Q, radii, lengths.size)
    flops = flops_curve*args.batch
    achieved = flops/(batched["eval_kernel_ms"]*1e-3)/1e12
    ref_flops = reference_flops_cylinder_fq(C2_Q, np.repeat(radii, lengths.size))*args.batch
    generic["roofline_tflops_reference_formula"] = ref_flops/(generic["eval_kernel_ms"]*1e-3)/1e12
    # counter figures of the same launch from the committed profile summary
    traffic = valu_per_launch = None
    tfile = os.path.join(ROOT, "profiles", "k_batch_sep_cylinder_c2_counters.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            prof = json.load(f)
        traffic = prof.get("hbm_bytes_per_launch")
        valu_per_launch = prof.get("valu_wave_instructions_per_launch")
    from sasmodels_amd.kernelhip import probe_fma
    fma_tflops = probe_fma()
    valu_issue = None
    if valu_per_launch:
        # wave64 FMA instructions/s of the probe = TFLOP/s * 1e12 / (2 * 64)
        rate = valu_per_launch/(batched["eval_kernel_ms"]*1e-3)
        valu_issue = {"valu_wave_instructions_per_launch": valu_per_launch,
                      "valu_wave_instructions_per_s": rate,
                      "probe_fma_wave_instructions_per_s": fma_tflops*1e12/128.0,
                      "issue_frac": rate/(fma_tflops*1e12/128.0)}

    # CPU leg: oracle timing on whole curves; its curves are the FP64 references
    baseline, acc = None, {"max": None, "median": None}
    if not args.no_cpu:
        jobs = [(d, v, 1e-5) for d, v in ccalls[:4]] + [(d, v, 0.0) for d, v in gcalls[:4]] + \
               [(d, v, 0.0) for d, v in calls0]
        baseline, refs = cpu_baseline(info, jobs, threads)
        cut["accuracy_vs_fp64"] = accuracy(cout[:4], refs[:4])
        generic["accuracy_vs_fp64"] = accuracy(gout[:4], refs[4:8])
        batched["accuracy_vs_fp64"] = acc = accuracy(out0, refs[8:])
        batched["accuracy_curves"] = len(refs) - 8

    line = {
        "metric": METRIC,
        "value": batched["curves_per_s"],
        "unit": "curves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batched["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cylinder 1D: 1000 log q in [1e-3,1], radius 20 / length 400 lognormal pd 0.1 "
                               "(35x35 mesh), sld 4/1, Gauss-76, cutoff 0; %d curves per batched launch"
                               % args.batch,
                   "mesh_points": int(npts), "nq": int(C2_Q.size), "quadrature": QUAD},
        "max_rel_err_vs_fp64": acc["max"],
        "median_rel_err_vs_fp64": acc["median"],
        "qpd_evals_per_s": batched["curves_per_s"]*C2_Q.size*npts,
        "roofline": {"bound": "valu", "kernel": "k_batch_sep", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved/PEAK_FP32_TFLOPS,
                     "flops_per_launch": flops, "flops_per_curve": flops_curve,
                     "avg_kernel_ms": batched["eval_kernel_ms"], "traffic": traffic,
                     "reference_formula_tflops": ref_flops/(batched["eval_kernel_ms"]*1e-3)/1e12,
                     "measured_fp32_fma_peak_tflops": fma_tflops, "frac_of_measured_peak": achieved/fma_tflops,
                     "valu_issue": valu_issue},
        "paths": {"batched": batched, "batched_cutoff": cut, "batched_generic": generic,
                  "plain": plain, "plain_generic": plain_generic},
    }
    if not args.no_surrogate:
        line["paths"]["surrogate"] = bench_surrogate()
    if not args.no_configs:
        line["other_configs"] = bench_configs(threads, cpu=not args.no_cpu)
    if baseline:
        line["cpu_baseline"] = baseline
        line["speedup_vs_cpu"] = line["value"]/baseline["value"]
    model.release()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
