"""
Benchmark of the I(q) path on one MI355X.

Workload (the cylinder configuration the headline metric is quoted on):
cylinder 1D, 1000 log-spaced q in [1e-3, 1] 1/A, radius 20 A and length
400 A each with lognormal polydispersity 0.1 (35 x 35 mesh points, 3 sigma),
sld 4 / solvent 1, orientation average by fixed Gauss-76, FP32 kernels with
compensated sums.  A step evaluates one full I(q) curve through the
reference-shaped call (details + values in, result out); every step uses
its own parameter set (radius and length jittered by step), all packed
before the timed region.

Printed: one JSON line with curves/s, the roofline of the dominant kernel
(counted FLOPs of the reference formulas over its measured device time)
and the CPU oracle timed on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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

C2_PARS = dict(radius=20.0, length=400.0, sld=4.0, sld_solvent=1.0, scale=1.0, background=1e-3,
               radius_pd=0.1, radius_pd_n=35, radius_pd_nsigma=3, radius_pd_type="lognormal",
               length_pd=0.1, length_pd_n=35, length_pd_nsigma=3, length_pd_type="lognormal")
C2_Q = np.logspace(-3, 0, 1000)
QUAD = 76


def reference_flops_cylinder_fq(q, radii, lengths, weights_mask=None):
    """
    FLOPs of one cylinder <F>,<F^2> pass as the reference formulas state it
    (cylinder.c:55-88 with sas_J1.c:107-153, polevl.c, kernel_header.c:177)
    and the pd-loop accumulation (kernel_iq.c:797-799), counting add, sub,
    mul, div, sqrt, sin and cos as 1 each, for the branches taken:
      per Gauss node: 21 + J1 (29 for |x| <= 5, 67 beyond)
      per (q, mesh point): 18
    Returns the total over all q and mesh points.
    """
    z, _ = np.polynomial.legendre.leggauss(QUAD)
    theta = z*np.pi/4 + np.pi/4
    s = np.sin(theta)
    R = np.asarray(radii, "d")
    total = 0.0
    for qi in q:
        x = qi*R[:, None]*s[None, :]                     # [mesh point, node]
        small = np.count_nonzero(np.abs(x) <= 5.0)
        large = x.size - small
        total += small*(21 + 29) + large*(21 + 67) + 18*R.size
    return total


def build_steps(steps, info, kernel):
    """Per-step (details, values) with the radius/length jittered by step."""
    from sasmodels_amd import direct_model
    from sasmodels_amd.details import make_kernel_args
    out = []
    for k in range(steps):
        pars = dict(C2_PARS, radius=C2_PARS["radius"]*(1 + 1e-4*k), length=C2_PARS["length"]*(1 - 1e-4*k))
        mesh = direct_model.get_mesh(info, pars, dim="1d")
        out.append(make_kernel_args(kernel, mesh)[:2])
    return out


def cpu_baseline(info, details, values, threads, budget_s=15.0):
    """Oracle (FP64 C, OpenMP over q) on a q subsample sized to ~budget_s."""
    from oracle import oracle
    oracle.set_threads(threads)
    probe = C2_Q[::100]
    t0 = time.perf_counter()
    oracle.raw_call("cylinder", 0, QUAD, probe, details.buffer, values)
    t_probe = max(time.perf_counter() - t0, 1e-3)
    per_q = t_probe/probe.size
    nq = int(min(C2_Q.size, max(10, budget_s/per_q)))
    sample = C2_Q[np.linspace(0, C2_Q.size - 1, nq).astype(int)]
    t0 = time.perf_counter()
    oracle.raw_call("cylinder", 0, QUAD, sample, details.buffer, values)
    dt = time.perf_counter() - t0
    curve_s = dt*C2_Q.size/sample.size
    return {"value": 1.0/curve_s, "unit": "curves/s", "cores": threads, "kind": "port",
            "sample": "oracle/sasoracle.c FP64, %d of the 1000 q points of one curve (full 35x35 mesh, "
                      "Gauss-76), %d OpenMP threads, %.1f s, scaled to a whole curve" % (sample.size, threads, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from oracle.oracle import OracleKernel
    from sasmodels_amd import core, generate

    info = core.load_model_info("cylinder")
    generate.set_integration_size(info, QUAD)
    model = core.build_model(info, platform="hip")
    kernel = model.make_kernel([C2_Q])

    steps = build_steps(args.warmup + args.steps, info, kernel)

    # accuracy of the timed configuration against the FP64 oracle
    det0, val0 = steps[0]
    I_gpu = kernel(det0, val0, 0.0, False)
    from oracle import oracle as _o
    _o.set_threads(min(16, len(os.sched_getaffinity(0))))
    I_ref = OracleKernel(info, [C2_Q])(det0, val0, 0.0, False)
    sel = I_ref > 1e-6*I_ref.max()
    rel = np.abs(I_gpu[sel] - I_ref[sel])/I_ref[sel]

    for det, val in steps[:args.warmup]:
        kernel(det, val, 0.0, False)
    kernel.set_timing(True)
    t0 = time.perf_counter()
    for det, val in steps[args.warmup:]:
        kernel(det, val, 0.0, False)
    dt = time.perf_counter() - t0
    eval_ms, ncalls = kernel.get_timing()
    kernel.set_timing(False)

    # roofline of the evaluation kernel (mesh identical in size every step)
    from sasmodels_amd import direct_model
    mesh = direct_model.get_mesh(info, C2_PARS, dim="1d")
    radii = np.repeat(mesh[4][1], len(mesh[5][1]))
    flops = reference_flops_cylinder_fq(C2_Q, radii, None)
    avg_ms = eval_ms/max(ncalls, 1)
    achieved = flops/(avg_ms*1e-3)/1e12

    line = {
        "metric": METRIC,
        "value": args.steps/dt,
        "unit": "curves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt/args.steps*1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cylinder 1D: 1000 log q in [1e-3,1], radius 20 / length 400 lognormal pd 0.1 "
                               "(35x35), sld 4/1, Gauss-76, plain HIP path (one curve per call)",
                   "path": "plain", "mesh_points": int(radii.size), "nq": int(C2_Q.size),
                   "quadrature": QUAD},
        "max_rel_err_vs_fp64": float(rel.max()),
        "median_rel_err_vs_fp64": float(np.median(rel)),
        "roofline": {"bound": "valu", "kernel": "k_fq1d<Cylinder,fixed>", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved/PEAK_FP32_TFLOPS,
                     "flops_per_launch": flops, "avg_kernel_ms": avg_ms, "traffic": None},
    }
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        line["cpu_baseline"] = cpu_baseline(info, det0, val0, threads)
        line["speedup_vs_cpu"] = line["value"]/line["cpu_baseline"]["value"]
    kernel.release()
    model.release()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
