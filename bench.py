"""
Benchmark of the I(q) path on one MI355X.

Headline workload (the cylinder configuration the metric is quoted on):
cylinder 1D, 1000 log-spaced q in [1e-3, 1] 1/A, radius 20 A and length
400 A each with lognormal polydispersity 0.1 (35 x 35 mesh points, 3
sigma), sld 4 / solvent 1, orientation average by fixed Gauss-76, FP32
kernels with compensated sums and FP64 weight sums.

Paths, one JSON line:
  batched    (value)  a step evaluates BATCH curves of the workload in one
                      persistent launch; every curve has its own radius and
                      length (jittered by step and curve), all packed
                      before the timed region
  plain               one curve per reference-shaped call
  surrogate           the trained MLP on 4096 curves x 512 q drawn from its
                      box, against the plain kernel on the same curves
roofline: reference-formula FLOPs of the batched evaluation kernel over its
device time (HIP events on the batch's stream).  cpu_baseline: the FP64
oracle on the host cores, on a bounded sample of the workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu]
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


def _kernel_like(info):
    return type("K", (), {"info": info, "dtype": np.dtype("d")})()


def c2_calls(info, n, offset=0):
    """(details, values) for n jittered copies of the workload."""
    from sasmodels_amd import direct_model
    from sasmodels_amd.details import make_kernel_args
    kl = _kernel_like(info)
    out = []
    for k in range(offset, offset + n):
        f = 1 + 1e-4*(k % 97)
        pars = dict(C2_PARS, radius=C2_PARS["radius"]*f, length=C2_PARS["length"]*(2 - f))
        out.append(make_kernel_args(kl, direct_model.get_mesh(info, pars, dim="1d"))[:2])
    return out


def masked_relerr(a, b):
    sel = b > 1e-6*b.max()
    return np.abs(a[sel] - b[sel])/b[sel]


def cpu_baseline(details, values, threads, budget_s=15.0):
    """Oracle (FP64 C, OpenMP over q) on a q subsample sized to ~budget_s."""
    from oracle import oracle
    oracle.set_threads(threads)
    probe = C2_Q[::100]
    t0 = time.perf_counter()
    oracle.raw_call("cylinder", 0, QUAD, probe, details.buffer, values)
    per_q = max(time.perf_counter() - t0, 1e-3)/probe.size
    nq = int(min(C2_Q.size, max(10, budget_s/per_q)))
    sample = C2_Q[np.linspace(0, C2_Q.size - 1, nq).astype(int)]
    t0 = time.perf_counter()
    oracle.raw_call("cylinder", 0, QUAD, sample, details.buffer, values)
    dt = time.perf_counter() - t0
    curve_s = dt*C2_Q.size/sample.size
    return {"value": 1.0/curve_s, "unit": "curves/s", "cores": threads, "kind": "port",
            "sample": "oracle/sasoracle.c FP64 on %d of the 1000 q points of one workload curve "
                      "(full 35x35 mesh, Gauss-76), %d OpenMP threads, %.1f s, scaled to a whole curve"
                      % (sample.size, threads, dt)}


def bench_plain(model, info, steps, warmup):
    kernel = model.make_kernel([C2_Q])
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


def bench_batched(model, info, steps, warmup, batch_size):
    from sasmodels_amd.kernelhip import HipBatch
    batch = HipBatch(model, [[C2_Q]]*batch_size, use_graph=True)
    groups = [c2_calls(info, batch_size, offset=g*batch_size) for g in range(steps + warmup)]
    pts, waves = batch.autotune(groups[0])
    for calls in groups[:warmup]:
        batch(calls)
    t0 = time.perf_counter()
    for calls in groups[warmup:]:
        out = batch(calls)
    dt = time.perf_counter() - t0
    # device time of the evaluation kernel (eager launches with events)
    batch.set_timing(True)
    for calls in groups[warmup:warmup + min(steps, 10)]:
        batch(calls)
    ms, n = batch.get_timing()
    batch.set_timing(False)
    first = batch(groups[0])
    batch.release()
    return ({"curves_per_s": steps*batch_size/dt, "ms_per_step": dt/steps*1e3, "batch": batch_size,
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
    ref = np.array(b(calls))
    b.set_timing(True)
    t0 = time.perf_counter()
    b(calls)
    t_plain = time.perf_counter() - t0
    pms, _ = b.get_timing()
    b.release()
    err = np.array([masked_relerr(p, r).max() for p, r in zip(pred, ref)])
    flops = 2.0*198400*ncurves*nq
    return {"curves_per_s": ncurves/(ms*1e-3), "ms_per_pass": ms, "curves": ncurves, "nq": nq,
            "tflops_bf16": flops/(ms*1e-3)/1e12,
            "max_rel_err_vs_plain": float(err.max()), "median_rel_err_vs_plain": float(np.median(err)),
            "frac_curves_within_1e-3": float(np.mean(err <= 1e-3)),
            "frac_curves_within_1e-2": float(np.mean(err <= 1e-2)),
            "plain_curves_per_s_same_curves": nsub/t_plain,
            "plain_eval_kernel_ms_per_curve": pms/nsub,
            "speedup_vs_plain": (ncurves/(ms*1e-3))/(nsub/t_plain)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a kernel of known traffic for "
                         "scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from oracle import oracle
    from oracle.oracle import OracleKernel
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
    oracle.set_threads(threads)
    info = core.load_model_info("cylinder")
    generate.set_integration_size(info, QUAD)
    model = core.build_model(info, platform="hip")

    batched, calls0, out0 = bench_batched(model, info, args.steps, args.warmup, args.batch)
    plain = bench_plain(model, info, args.steps, args.warmup)

    # accuracy of the timed workload against the FP64 oracle (first curve of the first batch)
    det0, val0 = calls0[0]
    I_ref = OracleKernel(info, [C2_Q])(det0, val0, 0.0, False)
    rel = masked_relerr(np.asarray(out0[0]), I_ref)

    mesh = direct_model.get_mesh(info, C2_PARS, dim="1d")
    radii = np.repeat(mesh[4][1], len(mesh[5][1]))
    flops_curve = reference_flops_cylinder_fq(C2_Q, radii)
    flops = flops_curve*args.batch
    achieved = flops/(batched["eval_kernel_ms"]*1e-3)/1e12

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
                               "(35x35 mesh), sld 4/1, Gauss-76; %d curves per batched launch" % args.batch,
                   "mesh_points": int(radii.size), "nq": int(C2_Q.size), "quadrature": QUAD},
        "max_rel_err_vs_fp64": float(rel.max()),
        "median_rel_err_vs_fp64": float(np.median(rel)),
        "roofline": {"bound": "valu", "kernel": "k_batch_eval_1d<Cylinder,fixed>", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved/PEAK_FP32_TFLOPS,
                     "flops_per_launch": flops, "flops_per_curve": flops_curve,
                     "avg_kernel_ms": batched["eval_kernel_ms"], "traffic": None},
        "paths": {"batched": batched, "plain": plain},
    }
    if not args.no_surrogate:
        line["paths"]["surrogate"] = bench_surrogate()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(det0, val0, threads)
        line["speedup_vs_cpu"] = line["value"]/line["cpu_baseline"]["value"]
    model.release()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
