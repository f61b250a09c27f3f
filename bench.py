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
                            its box with the plain kernel as judge (8 q per
                            curve) and fallback inside the timed work,
                            against the plain batched kernel on the same
                            curves (device-resident)
roofline (every configuration): the FLOPs of its launch counted with the
reference formulas (profiles/flops.json, written by oracle/count_flops.py
from the same workload builders) over the mean device time of its dominant
kernel (HIP events on the launch's stream), against the published FP32 peak
(BF16 dense MFMA for the surrogate's k_infer); traffic and the executed work
(VALU instructions issued, FLOPs the hardware counted) from the counter
summaries in profiles/ when present.  cpu_baseline: the FP64 oracle on the
host cores (every core of the affinity mask, capped by OMP_NUM_THREADS), on
a bounded sample of the workload.

other_configs: the sphere 1D, ellipsoid 2D, batched core-shell cylinder,
batched magnetic core-shell sphere, batched parallelepiped (nested
orientation rules), batched bcc paracrystal (FP64), pinhole-smeared cylinder
and cylinder@hardsphere (beta) configurations, each with device throughput,
FP64 accuracy on a subset and the oracle's rate on the host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu]
                    [--no-surrogate] [--no-configs] [--only LEG]
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
PEAK_BF16_TFLOPS = 2500.0
SEED = 20261203
SURROGATE_WEIGHTS = os.path.join(ROOT, "sasmodels_amd", "data", "cylinder_surrogate.npz")

C2_PARS = dict(radius=20.0, length=400.0, sld=4.0, sld_solvent=1.0, scale=1.0, background=1e-3,
               radius_pd=0.1, radius_pd_n=35, radius_pd_nsigma=3, radius_pd_type="lognormal",
               length_pd=0.1, length_pd_n=35, length_pd_nsigma=3, length_pd_type="lognormal")
C2_Q = np.logspace(-3, 0, 1000)
QUAD = 76


def _kernel_like(info):
    return type("K", (), {"info": info, "dtype": np.dtype("d")})()


# ---------------------------------------------------------------- workloads
# Builders of the synthetic inputs of every configuration (no GPU): the
# benchmark and the FLOP count (oracle/count_flops.py) use the same ones.
# Each returns a dict: name, info (model info with its quadrature set),
# kind (0 Iq, 1 Iqxy, 2 Imagnetic), quad, q_list ([q] or [qx, qy] per set),
# calls ([(details, values)] per set), workload (description).

def c2_calls(info, n, offset=0, extra=None):
    """(details, values) for n jittered copies of the headline curve (plus *extra* parameters)."""
    from sasmodels_amd import direct_model
    from sasmodels_amd.details import make_kernel_args
    kl = _kernel_like(info)
    out = []
    for k in range(offset, offset + n):
        f = 1 + 1e-4*(k % 97)
        pars = dict(C2_PARS, radius=C2_PARS["radius"]*f, length=C2_PARS["length"]*(2 - f), **(extra or {}))
        out.append(make_kernel_args(kl, direct_model.get_mesh(info, pars, dim="1d"))[:2])
    return out


def _info(name, quad=None):
    from sasmodels_amd import core, generate
    info = core.load_model_info(name)
    if quad is not None:
        generate.set_integration_size(info, quad)
    return info


def _pack_all(info, pars_list, dim):
    from sasmodels_amd import direct_model
    from sasmodels_amd.details import make_kernel_args
    kl = _kernel_like(info)
    return [make_kernel_args(kl, direct_model.get_mesh(info, p, dim=dim))[:2] for p in pars_list]


def wl_cylinder(n=64, offset=0, cutoff=0.0):
    info = _info("cylinder", QUAD)
    return dict(name="cylinder_1d", info=info, kind=0, quad=QUAD, q_list=[[C2_Q]]*n, cutoff=cutoff,
                calls=c2_calls(info, n, offset),
                workload="cylinder 1D: 1000 log q in [1e-3,1], radius 20 / length 400 lognormal pd 0.1 "
                         "(35x35 mesh), sld 4/1, Gauss-76, cutoff %g; %d curves per batched launch" % (cutoff, n))


def wl_cylinder_adaptive(n=64, offset=0):
    info = _info("cylinder", 0)
    return dict(name="cylinder_1d_adaptive", info=info, kind=0, quad=0, q_list=[[C2_Q]]*n, cutoff=0.0,
                calls=c2_calls(info, n, offset),
                workload="cylinder 1D as the headline curve (1000 log q in [1e-3,1], radius 20 / length 400 "
                         "lognormal pd 0.1, 35x35 mesh) with the reference's default adaptive orientation rule "
                         "(20/76/500/5000 points by q*max(R, L/2)); %d curves per batched launch" % n)


def wl_sphere(n=4096):
    info = _info("sphere")
    q = np.logspace(-3, np.log10(0.5), 200)
    pars = [dict(radius=50*(1 + 1e-3*(k % 101)), radius_pd=0.1, radius_pd_n=35, radius_pd_nsigma=3, sld=4,
                 sld_solvent=1, scale=1, background=1e-3) for k in range(n)]
    return dict(name="sphere_1d", info=info, kind=0, quad=0, q_list=[[q]]*n, cutoff=0.0,
                calls=_pack_all(info, pars, "1d"),
                workload="sphere 1D, 200 q in [1e-3, 0.5], radius 50 gaussian pd 0.1 (35 points), "
                         "%d curves per launch" % n)


def wl_ellipsoid_2d():
    info = _info("ellipsoid")
    x = np.linspace(-0.3, 0.3, 256)
    qx, qy = [a.ravel() for a in np.meshgrid(x, x)]
    pars = dict(radius_polar=20, radius_equatorial=400, theta=60, phi=60, theta_pd=10, theta_pd_n=20,
                phi_pd=20, phi_pd_n=20, radius_equatorial_pd=0.1, radius_equatorial_pd_n=20)
    return dict(name="ellipsoid_2d", info=info, kind=1, quad=0, q_list=[[qx, qy]], cutoff=0.0,
                calls=_pack_all(info, [pars], "2d"),
                workload="ellipsoid 2D, 256 x 256 detector over [-0.3, 0.3]^2, radius_polar 20, "
                         "radius_equatorial 400 (gaussian pd 0.1, 20 points), theta 60 / phi 60 with "
                         "gaussian jitter 10 / 20 deg (20 points each): 8000-point mesh")


def wl_core_shell_cylinder(n=4096):
    info = _info("core_shell_cylinder", QUAD)
    rng = np.random.default_rng(SEED + 4)
    q_list, pars = [], []
    for _ in range(n):
        nq = int(rng.integers(64, 2049))
        q_list.append([np.logspace(-3, np.log10(0.5), nq)])
        pars.append(dict(radius=rng.uniform(10, 100), thickness=rng.uniform(5, 30), length=rng.uniform(50, 1000),
                         sld_core=rng.uniform(-1, 7), sld_shell=rng.uniform(-1, 7), sld_solvent=rng.uniform(-1, 7),
                         radius_pd=rng.uniform(0, 0.2), radius_pd_n=int(rng.integers(1, 36))))
    return dict(name="core_shell_cylinder_batched", info=info, kind=0, quad=QUAD, q_list=q_list, cutoff=0.0,
                calls=_pack_all(info, pars, "1d"),
                workload="%d core_shell_cylinder sets, nq U{64..2048} log-spaced in [1e-3, 0.5], radius U[10,100], "
                         "thickness U[5,30], length U[50,1000], slds U[-1,7], radius pd U[0,0.2] x U{1..35} "
                         "gaussian points, Gauss-76, one persistent launch" % n)


def wl_magnetic_core_shell_sphere(n=256):
    info = _info("core_shell_sphere")
    rng = np.random.default_rng(SEED + 5)
    x = np.linspace(-0.3, 0.3, 128)
    qx, qy = [a.ravel() for a in np.meshgrid(x, x)]
    pars = [dict(sld_core_M0=rng.uniform(0, 3), sld_shell_M0=rng.uniform(0, 3), sld_solvent_M0=rng.uniform(0, 3),
                 sld_core_mtheta=rng.uniform(0, 90), sld_core_mphi=rng.uniform(0, 360),
                 sld_shell_mtheta=rng.uniform(0, 90), sld_shell_mphi=rng.uniform(0, 360),
                 sld_solvent_mtheta=rng.uniform(0, 90), sld_solvent_mphi=rng.uniform(0, 360),
                 up_frac_i=rng.uniform(0, 1), up_frac_f=rng.uniform(0, 1), up_theta=rng.uniform(0, 90))
            for _ in range(n)]
    return dict(name="magnetic_core_shell_sphere_batched", info=info, kind=2, quad=0, q_list=[[qx, qy]]*n,
                cutoff=0.0, calls=_pack_all(info, pars, "2d"),
                workload="%d magnetic core_shell_sphere sets on a 128 x 128 detector over [-0.3, 0.3]^2, "
                         "sld_*_M0 U[0,3], mtheta U[0,90], mphi U[0,360], up_frac_i/f U[0,1], up_theta U[0,90]: "
                         "four spin states" % n)


def wl_parallelepiped(n=256):
    info = _info("parallelepiped")
    rng = np.random.default_rng(SEED + 6)
    q = np.logspace(-3, np.log10(0.5), 200)
    pars = [dict(length_a=rng.uniform(20, 60), length_b=rng.uniform(40, 120), length_c=rng.uniform(100, 400),
                 length_a_pd=0.1, length_a_pd_n=5, length_b_pd=0.1, length_b_pd_n=5) for _ in range(n)]
    return dict(name="parallelepiped_batched", info=info, kind=0, quad=0, q_list=[[q]]*n, cutoff=0.0,
                calls=_pack_all(info, pars, "1d"),
                workload="%d parallelepiped sets, 200 q in [1e-3, 0.5], adaptive nested (theta, phi) rules, "
                         "5 x 5 length_a x length_b mesh, one persistent launch" % n)


def wl_bcc(n=64):
    info = _info("bcc_paracrystal")
    rng = np.random.default_rng(SEED + 7)
    q = np.logspace(-3, np.log10(0.5), 200)
    pars = [dict(dnn=rng.uniform(150, 300), d_factor=rng.uniform(0.05, 0.2), radius=rng.uniform(20, 50),
                 radius_pd=0.1, radius_pd_n=5) for _ in range(n)]
    return dict(name="bcc_paracrystal_batched", info=info, kind=0, quad=0, q_list=[[q]]*n, cutoff=0.0,
                calls=_pack_all(info, pars, "1d"),
                workload="%d bcc_paracrystal sets, 200 q in [1e-3, 0.5], FP64, Gauss-150 x 150 (theta, phi), "
                         "5-point radius mesh, one persistent launch" % n)


#: configurations whose FLOPs oracle/count_flops.py counts (name -> builder)
WORKLOADS = {
    "cylinder_1d": wl_cylinder,
    "cylinder_1d_cutoff": lambda: wl_cylinder(cutoff=1e-5),
    "cylinder_1d_adaptive": wl_cylinder_adaptive,
    "sphere_1d": wl_sphere,
    "ellipsoid_2d": wl_ellipsoid_2d,
    "core_shell_cylinder_batched": wl_core_shell_cylinder,
    "magnetic_core_shell_sphere_batched": wl_magnetic_core_shell_sphere,
    "parallelepiped_batched": wl_parallelepiped,
    "bcc_paracrystal_batched": wl_bcc,
}

FLOPS_FILE = os.path.join(ROOT, "profiles", "flops.json")


def counted_flops():
    """Counted FLOPs per launch of each configuration (profiles/flops.json), or {}."""
    if not os.path.exists(FLOPS_FILE):
        return {}
    with open(FLOPS_FILE) as f:
        return json.load(f)


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


def host_threads():
    """(threads used by the CPU legs, threads the process may run on, nproc):
    every core of the affinity mask, capped by OMP_NUM_THREADS where the
    host sets one (a shared host may give each GPU a share of its cores)."""
    affinity = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(affinity, int(omp)) if omp.isdigit() and int(omp) > 0 else affinity
    return threads, affinity, os.cpu_count()


def cpu_baseline(info, jobs, threads, budget_s=12.0):
    """
    The CPU leg: the FP64 oracle (C, OpenMP over q) on whole workload curves
    -- *jobs* is a list of (details, values, cutoff) -- in order until
    budget_s of host time is spent, then one whole curve on one thread.
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
    det, val, _ = jobs[-1]
    t1 = time.perf_counter()
    ok(det, val, 0.0, False)
    dt1 = time.perf_counter() - t1
    oracle.set_threads(threads)
    _, affinity, nproc = host_threads()
    rec = {"value": len(refs)/dt, "unit": "curves/s", "cores": threads, "kind": "port",
           "sample": "oracle/sasoracle.c FP64 (the reference formulas, OpenMP over q), %d whole workload curves "
                     "(1000 q, 35x35 mesh, Gauss-76) in %.1f s on %d threads" % (len(refs), dt, threads),
           "host_affinity_cores": affinity, "host_nproc": nproc,
           "single_thread_curves_per_s": 1.0/dt1,
           "single_thread_sample": "one whole workload curve on 1 thread, %.2f s" % dt1}
    return rec, refs


def bench_plain(model, info, steps, warmup, separable=True):
    kernel = model.make_kernel([C2_Q])
    kernel.set_separable(separable)
    calls = c2_calls(info, steps + warmup, offset=1000)
    for det, val in calls[:warmup]:
        kernel(det, val, 0.0, False)
    t0 = time.perf_counter()
    for det, val in calls[warmup:]:
        kernel(det, val, 0.0, False)
    dt = time.perf_counter() - t0
    # the evaluation kernel's time in a pass of its own (the events would
    # add to the calls above)
    kernel.set_timing(True)
    for det, val in calls[warmup:warmup + min(steps, 20)]:
        kernel(det, val, 0.0, False)
    ms, n = kernel.get_timing()
    kernel.release()
    return {"curves_per_s": steps/dt, "ms_per_curve": dt/steps*1e3, "eval_kernel_ms": ms/max(n, 1)}


def kernel_ms(batch, calls, reps=10, **kw):
    """Mean device time of the batch's evaluation kernel over *reps* eager
    calls on *calls* (HIP events on the batch's stream).  The calls write
    I(q) to an ordinary host array, so the kernels' outputs stay in HBM (a
    pinned output array would have the evaluation kernel of single-chunk
    sets write across PCIe)."""
    n = int(batch.qoff[-1]) if batch._n_out is None else int(batch._n_out.sum())
    out = np.empty(n)
    batch.set_timing(True)
    for _ in range(reps):
        batch(calls, out=out, **kw)
    ms, n = batch.get_timing()
    batch.set_timing(False)
    return ms/max(n, 1)


def resident_s(batch, reps):
    """Seconds per device-resident run of the last call's device work."""
    batch.replay(2)
    t0 = time.perf_counter()
    batch.replay(reps)
    return (time.perf_counter() - t0)/reps


def bench_batched(model, info, steps, warmup, batch_size, separable=True, cutoff=0.0):
    """
    Two timings of a batch of the workload: (1) device-resident, the value:
    `steps` back-to-back runs of the device work (expand, evaluate, final
    sums) with inputs and outputs in HBM, wall clock bracketed by
    synchronisation; (2) host-inclusive: `steps` full calls with different
    jittered inputs each (host packing, staging upload, hipGraph replay,
    result download).  The evaluation kernel is timed with HIP events on the
    inputs whose FLOPs profiles/flops.json counts (the first group).
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
        batch(calls, cutoff)
    dt_host = time.perf_counter() - t0
    ms = kernel_ms(batch, groups[0], min(steps, 10), cutoff=cutoff)
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
             "eval_kernel_ms": ms}, groups[0], first)


# ---------------------------------------------------------------- rooflines
def profile_counters(stem):
    """Counter summary profiles/<stem>_counters.json, or {}."""
    path = os.path.join(ROOT, "profiles", stem + "_counters.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def roofline(config, kernel, ms, sets=None, stem=None, fma_tflops=None):
    """
    Roofline record of the dominant kernel of *config*: achieved = the
    counted FLOPs of its launch (profiles/flops.json, the reference formulas
    on the counting oracle; scaled by sets/counted sets when the launch has
    another number of sets) over its mean device time *ms*, against the
    published FP32 peak.  With a counter summary (profiles/<stem>_counters.json):
    traffic = HBM bytes per launch, and the executed work -- VALU
    instructions issued per second against the FMA issue rate the device
    sustains (saship_probe_fma), and the FP32/FP64 FLOPs the hardware
    counted -- beside the algorithmic figure.
    """
    rec = counted_flops().get(config)
    if rec is None or not ms:
        return None
    flops = rec["flops_per_launch"]*(1.0 if sets is None else sets/rec["sets"])
    achieved = flops/(ms*1e-3)/1e12
    out = {"bound": "valu", "kernel": kernel, "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
           "unit": "TFLOP/s", "frac": achieved/PEAK_FP32_TFLOPS, "traffic": None,
           "flops_per_launch": flops, "flops_per_qpd_eval": rec["flops_per_qpd_eval"],
           "avg_kernel_ms": ms, "flops_source": "profiles/flops.json (oracle/count_flops.py)"}
    prof = profile_counters(stem) if stem else {}
    if prof:
        out["traffic"] = prof.get("hbm_bytes_per_launch")
        out["counters"] = "profiles/%s_counters.json" % stem
        valu = prof.get("valu_wave_instructions_per_launch")
        if valu and fma_tflops:
            rate = valu/(ms*1e-3)
            probe = fma_tflops*1e12/128.0        # wave64 FMA instructions per second
            ex = {"valu_wave_instructions_per_launch": valu, "valu_issue_frac": rate/probe}
            hw = (prof.get("valu_flops_fp32_per_launch") or 0) + (prof.get("valu_flops_fp64_per_launch") or 0)
            if hw:
                ex["hw_counted_flops_per_launch"] = hw
                ex["hw_counted_tflops"] = hw/(ms*1e-3)/1e12
            out["executed"] = ex
    return out


# ---------------------------------------------------------------- the surrogate
def bench_surrogate(steps=20, ncurves=4096, nq=512, threshold=1e-3, judge_points=8):
    """
    The surrogate on `ncurves` curves x `nq` q drawn from its box, with the
    plain kernel as judge and fallback inside the timed work: a gated pass
    is inference (fused kernel), the plain kernel at `judge_points` q of
    every curve, and the plain kernel over the whole curve for every curve
    the judge rejects.  Times are device-resident (inputs in HBM); the
    accuracy is measured against the plain kernel on every curve.
    """
    if not os.path.exists(SURROGATE_WEIGHTS):
        return None
    import torch
    from sasmodels_amd.surrogate import CylinderSurrogate, in_box, sample_box
    sur = CylinderSurrogate()
    sur.load(SURROGATE_WEIGHTS)
    rng = np.random.default_rng(SEED)
    curves = sample_box(ncurves, rng)
    lq = np.linspace(-3, 0, nq)
    q = 10.0**lq
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
    t_inf = e0.elapsed_time(e1)/steps*1e-3
    pred = out.view(ncurves, nq).double().cpu().numpy()
    calls = sur.plain.calls(curves)
    # plain kernel over whole curves: the reference of the accuracy and the fallback's cost
    b = sur.plain.batch([q]*ncurves)
    b.autotune(calls)
    ref = np.array(b(calls))
    t_plain = resident_s(b, 5)
    b.release()
    # judge: the plain kernel at judge_points q of every curve
    jidx = np.unique(np.linspace(0, nq - 1, judge_points).round().astype(int))
    bj = sur.plain.batch([q[jidx]]*ncurves)
    judge = np.array(bj(calls))
    t_judge = resident_s(bj, 5)
    bj.release()
    err = np.array([masked_relerr(p, r).max() for p, r in zip(pred, ref)])
    jerr = np.max(np.abs(pred[:, jidx] - judge)/np.abs(judge), axis=1)
    reject = (jerr > threshold) | ~in_box(curves, lq)
    final_err = np.where(reject, 0.0, err)
    t_gated = t_inf + t_judge + reject.mean()*t_plain
    flops = 2.0*198400*ncurves*nq
    return {"workload": "%d curves x %d log q in [1e-3, 1] uniform in the surrogate's box "
                        "(radius [10,200], length [20,2000], lognormal pd [0,0.3] x 35 points each, "
                        "contrast [0.1,8]); plain kernel Gauss-%d" % (ncurves, nq, QUAD),
            "curves_per_s_gated": ncurves/t_gated, "ms_per_pass_gated": t_gated*1e3,
            "curves_per_s_inference_only": ncurves/t_inf, "ms_inference": t_inf*1e3,
            "ms_judge": t_judge*1e3, "ms_plain_whole_batch": t_plain*1e3,
            "fallback_frac": float(reject.mean()), "judge_points": int(jidx.size), "threshold": threshold,
            "served_at_threshold": bool(sur.serves(threshold)),
            "policy": "CylinderSurrogate.evaluate serves threshold %g with these weights" % threshold
                      if sur.serves(threshold) else
                      "CylinderSurrogate.evaluate sends every request at threshold %g to the plain kernel "
                      "(validation record below %g of curves): the path runs at the plain rate; the gated "
                      "figures are what serving would cost" % (threshold, 0.5),
            "plain_curves_per_s_same_curves": ncurves/t_plain,
            "speedup_vs_plain_gated": t_plain/t_gated,
            "frac_curves_within_1e-3_ungated": float(np.mean(err <= 1e-3)),
            "frac_curves_within_1e-2_ungated": float(np.mean(err <= 1e-2)),
            "median_rel_err_vs_plain_ungated": float(np.median(err)),
            "max_rel_err_vs_plain_gated": float(final_err.max()),
            "escaped_frac": float(np.mean(~reject & (err > threshold))),
            "roofline": {"bound": "mfma", "kernel": "k_infer", "achieved": flops/t_inf/1e12,
                         "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                         "frac": flops/t_inf/1e12/PEAK_BF16_TFLOPS, "traffic": None,
                         "flops_per_launch": flops, "avg_kernel_ms": t_inf*1e3}}


# ---------------------------------------------------------------- other configurations
def _timed(fn, reps):
    fn()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    return (time.perf_counter() - t0)/reps, out


def _oracle_refs(wl, idx, threads):
    """Oracle I(q) of sets idx of workload wl and the seconds it took."""
    from oracle import oracle
    from oracle.oracle import OracleKernel
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    refs = [OracleKernel(wl["info"], wl["q_list"][i])(*wl["calls"][i], wl["cutoff"], wl["kind"] == 2)
            for i in idx]
    return time.perf_counter() - t0, refs


def bench_workload(wl, kernel_name, stem, fma_tflops, threads, cpu=True, reps=10, cpu_sets=4, single_sets=64):
    """
    One batched configuration: device-resident launches per second (replays
    of the staged call), host-inclusive call rate, the evaluation kernel's
    mean time (HIP events), the roofline, the rate of the first
    `single_sets` sets called one kernel call each (and the batch's
    resident rate over it), the FP64 accuracy of `cpu_sets` sets and the
    oracle's rate on them (the CPU leg, scaled by mesh work).
    """
    from sasmodels_amd import core
    from sasmodels_amd.kernelhip import HipBatch
    model = core.build_model(wl["info"], platform="hip")
    b = HipBatch(model, wl["q_list"], use_graph=True)
    mag = wl["kind"] == 2
    calls = wl["calls"]
    b.autotune(calls, wl["cutoff"], magnetic=mag)
    dt, iq = _timed(lambda: b(calls, wl["cutoff"], magnetic=mag), max(2, reps//3))
    buf = np.empty(int(b.qoff[-1]))
    dt_out, _ = _timed(lambda: b(calls, wl["cutoff"], magnetic=mag, out=buf), max(2, reps//3))
    ms = kernel_ms(b, calls, max(2, reps//3), cutoff=wl["cutoff"], magnetic=mag)
    b(calls, wl["cutoff"], magnetic=mag, out=buf)     # staged for the replays, outputs in HBM
    res = resident_s(b, reps)
    b.release()
    # the same sets one call each (the loop a caller without batches runs):
    # a kernel per set made beforehand, each call timed host-inclusive
    idx1 = range(min(len(calls), single_sets))
    kernels = [model.make_kernel(wl["q_list"][i]) for i in idx1]
    for k, i in zip(kernels, idx1):
        k(*calls[i], wl["cutoff"], mag)
    t0 = time.perf_counter()
    for k, i in zip(kernels, idx1):
        k(*calls[i], wl["cutoff"], mag)
    dt1 = (time.perf_counter() - t0)/len(kernels)
    for k in kernels:
        k.release()
    model.release()
    n = len(calls)
    work = [np.asarray(qv[0]).size*int(d.num_eval) for qv, (d, _) in zip(wl["q_list"], calls)]
    ent = {"workload": wl["workload"], "sets_per_s": n/res, "sets_per_s_host_inclusive": n/dt,
           "sets_per_s_host_inclusive_reused_output": n/dt_out,
           "ms_per_launch": res*1e3, "eval_kernel_ms": ms, "qpd_evals_per_s": sum(work)/(ms*1e-3),
           "single_call_sets_per_s": 1.0/dt1, "single_call_sets": len(kernels),
           "batch_over_single_calls": (n/res)*dt1,
           "roofline": roofline(wl["name"], kernel_name, ms, stem=stem, fma_tflops=fma_tflops)}
    if cpu:
        idx = list(range(min(cpu_sets, n)))
        t, refs = _oracle_refs(wl, idx, threads)
        ent["max_rel_err_vs_fp64"] = float(max(masked_relerr(np.asarray(iq[i]), r).max()
                                                for i, r in zip(idx, refs)))
        ent["accuracy_sets"] = len(idx)
        ent["cpu_sets_per_s"] = n/(t*sum(work)/sum(work[i] for i in idx))
        ent["cpu_threads"] = threads
    return ent


def bench_ellipsoid_2d(fma_tflops, threads, cpu=True, reps=5):
    """The 2D detector configuration: one image per reference-shaped call."""
    from sasmodels_amd import core
    wl = wl_ellipsoid_2d()
    qx, qy = wl["q_list"][0]
    det, val = wl["calls"][0]
    model = core.build_model(wl["info"], platform="hip")
    k = model.make_kernel([qx, qy])
    k(det, val, 0.0, False)
    k.set_timing(True)
    dt, img = _timed(lambda: k(det, val, 0.0, False), reps)
    ms, n = k.get_timing()
    k.release()
    model.release()
    ms /= max(n, 1)
    ent = {"workload": wl["workload"], "images_per_s": 1.0/dt, "eval_kernel_ms": ms,
           "qpd_evals_per_s": qx.size*int(det.num_eval)/(ms*1e-3),
           "roofline": roofline("ellipsoid_2d", "k_iq2d<Ellipsoid>", ms, stem="k_iq2d_ellipsoid_2d",
                                fma_tflops=fma_tflops)}
    if cpu:
        sub = np.random.default_rng(SEED).choice(qx.size, 1024, replace=False)
        from oracle import oracle
        from oracle.oracle import OracleKernel
        oracle.set_threads(threads)
        t0 = time.perf_counter()
        ref = OracleKernel(wl["info"], [qx[sub], qy[sub]])(det, val, 0.0, False)
        t = time.perf_counter() - t0
        ent["max_rel_err_vs_fp64"] = float(masked_relerr(np.asarray(img)[sub], ref).max())
        ent["max_rel_err_pixels"] = 1024
        ent["cpu_images_per_s"] = 1024/(t*qx.size)
        ent["cpu_threads"] = threads
    return ent


CONFIG_LEGS = ("sphere_1d", "ellipsoid_2d", "core_shell_cylinder_batched", "magnetic_core_shell_sphere_batched",
               "parallelepiped_batched", "bcc_paracrystal_batched", "cylinder_batched_pinhole",
               "cylinder_hardsphere_beta_batched", "cylinder_adaptive_batched")


def bench_configs(fma_tflops, threads, cpu=True, only=None):
    """The other configurations, one entry each."""
    from sasmodels_amd import core, generate
    from sasmodels_amd.kernelhip import HipBatch
    out = {}
    want = (lambda name: only is None or name == only)
    if want("sphere_1d"):
        out["sphere_1d"] = bench_workload(wl_sphere(), "k_batch_eval_1d<Sphere>", "k_batch_eval_1d_sphere_batched",
                                          fma_tflops, threads, cpu, cpu_sets=64)
    if want("ellipsoid_2d"):
        out["ellipsoid_2d"] = bench_ellipsoid_2d(fma_tflops, threads, cpu)
    if want("core_shell_cylinder_batched"):
        out["core_shell_cylinder_batched"] = bench_workload(
            wl_core_shell_cylinder(), "k_batch_eval_1d<CoreShellCylinder>",
            "k_batch_eval_1d_core_shell_cylinder_batched", fma_tflops, threads, cpu, reps=3, cpu_sets=16)
    if want("magnetic_core_shell_sphere_batched"):
        out["magnetic_core_shell_sphere_batched"] = bench_workload(
            wl_magnetic_core_shell_sphere(), "k_batch_eval_2d<CoreShellSphere, magnetic>",
            "k_batch_eval_2d_magnetic_core_shell_sphere", fma_tflops, threads, cpu, reps=3, cpu_sets=2)
    if want("parallelepiped_batched"):
        out["parallelepiped_batched"] = bench_workload(
            wl_parallelepiped(), "k_batch_eval_1d<Parallelepiped>", None, fma_tflops, threads, cpu, reps=5)
    if want("bcc_paracrystal_batched"):
        out["bcc_paracrystal_batched"] = bench_workload(
            wl_bcc(), "k_batch_eval_1d<BccParacrystal> (FP64)", None, fma_tflops, threads, cpu, reps=3)
    if want("cylinder_adaptive_batched"):
        out["cylinder_adaptive_batched"] = bench_workload(
            wl_cylinder_adaptive(), "k_batch_sep_adapt", "k_batch_sep_adapt_cylinder", fma_tflops, threads, cpu,
            reps=10, cpu_sets=2)

    if want("cylinder_batched_pinhole"):
        # headline curves measured through a pinhole (dq/q = 5 %): the batch
        # evaluates q_calc and smears on the device
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
        with_smear = resident_s(b, 20)
        b.set_resolution(None)
        b(calls)
        without = resident_s(b, 20)
        b.release()
        _, length, _, _ = resolution.band(res.weight_matrix)
        smear_bytes = 64*(16*int(length.sum()) + 8*C2_Q.size)
        ent = {"workload": "cylinder headline curves (64 per launch) with pinhole resolution dq/q = 0.05: "
                           "%d q_calc points smeared to 1000 q on the device" % res.q_calc.size,
               "curves_per_s": 64/with_smear, "curves_per_s_host_inclusive": 64/dt,
               "ms_per_launch": with_smear*1e3, "smear_ms": (with_smear - without)*1e3,
               "smear_gbytes_per_s": smear_bytes/max(with_smear - without, 1e-9)/1e9}
        if cpu:
            wlp = dict(info=info, q_list=[[res.q_calc]]*4, calls=calls[:4], cutoff=0.0, kind=0)
            t, refs = _oracle_refs(wlp, range(4), threads)
            ent["max_rel_err_vs_fp64"] = float(max(masked_relerr(np.asarray(a), res.apply(r)).max()
                                                    for a, r in zip(iq, refs)))
        out["cylinder_batched_pinhole"] = ent
        model.release()

    if want("cylinder_hardsphere_beta_batched"):
        # the headline curves as cylinder@hardsphere, beta approximation,
        # excluded-volume effective radius: S(q) and the combination fused
        # into the batch's final kernel (FP64)
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
        with_s = resident_s(pb, 20)
        p_calls = pb._split(calls)[0]
        pb.batch.set_structure(None)
        pb.batch(p_calls, radius_effective_mode=1)
        without = resident_s(pb, 20)
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
    ap.add_argument("--only", choices=("headline", "surrogate") + CONFIG_LEGS,
                    help="run one leg only and print its record (for profiling)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a kernel of known traffic for "
                         "scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from sasmodels_amd import core, direct_model, generate
    from sasmodels_amd.kernelhip import probe_fma

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy*(1 << 20)//4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    threads = host_threads()[0]
    fma_tflops = probe_fma()
    if args.only and args.only != "headline":
        rec = bench_surrogate(steps=5) if args.only == "surrogate" else \
            bench_configs(fma_tflops, threads, cpu=not args.no_cpu, only=args.only)[args.only]
        print(json.dumps({args.only: rec}))
        return

    info = core.load_model_info("cylinder")
    generate.set_integration_size(info, QUAD)
    model = core.build_model(info, platform="hip")

    batched, calls0, out0 = bench_batched(model, info, args.steps, args.warmup, args.batch)
    mesh = direct_model.get_mesh(info, C2_PARS, dim="1d")
    npts = mesh[4][1].size*mesh[5][1].size
    line_roof = roofline("cylinder_1d", "k_batch_sep", batched["eval_kernel_ms"], sets=args.batch,
                         stem="k_batch_sep_cylinder_c2", fma_tflops=fma_tflops)
    if line_roof:
        line_roof["note"] = ("achieved counts the reference algorithm's FLOPs (per q, mesh point and Gauss "
                             "node); the separable evaluation factorises radius x length and skips most of "
                             "them, so frac exceeds 1 -- 'executed' gives the kernel's own issue rate")
    if args.only == "headline":
        print(json.dumps({"batched": batched, "roofline": line_roof}))
        return
    short = max(2, min(args.steps, 10))
    cut, ccalls, cout = bench_batched(model, info, args.steps, args.warmup, args.batch, cutoff=1e-5)
    cut["roofline"] = roofline("cylinder_1d_cutoff", "k_batch_sep", cut["eval_kernel_ms"], sets=args.batch,
                               stem="k_batch_sep_cylinder_c2_cutoff", fma_tflops=fma_tflops)
    generic, gcalls, gout = bench_batched(model, info, short, 1, args.batch, separable=False)
    generic["roofline"] = roofline("cylinder_1d", "k_batch_eval_1d<Cylinder>", generic["eval_kernel_ms"],
                                   sets=args.batch, fma_tflops=fma_tflops)
    plain = bench_plain(model, info, args.steps, args.warmup)
    plain_generic = bench_plain(model, info, short, 1, separable=False)

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
        "roofline": line_roof,
        "measured_fp32_fma_tflops": fma_tflops,
        "paths": {"batched": batched, "batched_cutoff": cut, "batched_generic": generic,
                  "plain": plain, "plain_generic": plain_generic},
    }
    if not args.no_surrogate:
        line["paths"]["surrogate"] = bench_surrogate()
    if not args.no_configs:
        line["other_configs"] = bench_configs(fma_tflops, threads, cpu=not args.no_cpu)
    if baseline:
        line["cpu_baseline"] = baseline
        line["speedup_vs_cpu"] = line["value"]/baseline["value"]
    model.release()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
