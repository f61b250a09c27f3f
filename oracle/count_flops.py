"""
Algorithmic FLOPs of the benchmark configurations, counted with the
reference formulas: every set of a configuration's launch goes through the
counting oracle (libflopcount.so: sasoracle.c on a counting scalar), and the
tallies go to profiles/flops.json, which bench.py divides by kernel times.

TEST AND MEASUREMENT INFRASTRUCTURE: reads the workload builders of bench.py
(the same synthetic inputs the benchmark runs) and runs only the oracle.

    python -m oracle.count_flops [name ...] [--threads N]

Convention (oracle/flopcount.cpp): add, sub, mul, div, sqrt and each
transcendental (sin, cos, exp, log, pow, cbrt, asin, expm1) count 1; sign
changes, fabs, fmin/fmax and comparisons 0; only the branches taken count.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402

OUT = os.path.join(ROOT, "profiles", "flops.json")


def count_workload(wl, threads):
    """Summed tallies over every set of workload *wl* (a bench.py builder's dict)."""
    totals = dict.fromkeys(oracle.FLOP_CATEGORIES, 0)
    evals = 0
    for qv, (det, val) in zip(wl["q_list"], wl["calls"]):
        q = np.asarray(qv[0], "d") if wl["kind"] == 0 else np.column_stack([np.ravel(v) for v in qv])
        c = oracle.count_flops(wl["info"].id, wl["kind"], wl["quad"], q, det, val, wl["cutoff"], 0, threads)
        for k in oracle.FLOP_CATEGORIES:
            totals[k] += c[k]
        evals += q.shape[0]*int(det.num_eval)
    flops = sum(totals.values())
    return {"workload": wl["workload"], "sets": len(wl["calls"]), "flops_per_launch": flops,
            "qpd_evals_per_launch": evals, "flops_per_qpd_eval": flops/max(evals, 1),
            "categories": totals,
            "counted_with": "oracle/libflopcount.so: oracle/sasoracle.c (the reference formulas) on a "
                            "counting scalar, whole launch, branches as taken"}


def main(argv=None):
    import bench
    ap = argparse.ArgumentParser()
    ap.add_argument("names", nargs="*", help="configurations (default: all of bench.WORKLOADS)")
    ap.add_argument("--threads", type=int, default=len(os.sched_getaffinity(0)))
    args = ap.parse_args(argv)
    names = args.names or list(bench.WORKLOADS)
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for name in names:
        t0 = time.time()
        rec = count_workload(bench.WORKLOADS[name](), args.threads)
        rec["count_seconds"] = round(time.time() - t0, 1)
        data[name] = rec
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        print("%s: %.4g FLOPs per launch (%.1f per (q, mesh point)), %.0f s"
              % (name, rec["flops_per_launch"], rec["flops_per_qpd_eval"], rec["count_seconds"]), flush=True)


if __name__ == "__main__":
    main()
