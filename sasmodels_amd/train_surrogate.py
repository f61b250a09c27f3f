"""
Train the cylinder surrogate on one GPU and save its weights.

    python -m sasmodels_amd.train_surrogate [--curves 4096] [--nq 512]
        [--epochs 60] [--batch 8192] [--out sasmodels_amd/data/cylinder_surrogate.npz]

Training data: curves uniform in the surrogate's box, I(q) at nq log-spaced
q in [1e-3, 1] from the batched plain kernel (2M points by default).
Validation: held-out curves, relative error of the surrogate against the
plain kernel where I > 1e-6 max I, per curve.
"""

import argparse
import json
import os
import time

import numpy as np

from .surrogate import CylinderSurrogate, features, target

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "cylinder_surrogate.npz")


def curve_errors(pred, ref, mask=1e-6):
    """Per-curve max relative error over points with ref > mask * max ref."""
    err = np.abs(pred - ref)/np.abs(ref)
    sel = ref > mask*ref.max(axis=1, keepdims=True)
    return np.where(sel, err, 0.0).max(axis=1)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--curves", type=int, default=4096)
    ap.add_argument("--nq", type=int, default=512)
    ap.add_argument("--val", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args(argv)

    torch = __import__("torch")
    sur = CylinderSurrogate(seed=args.seed)
    t0 = time.time()
    curves, lq, Iq = sur.generate(args.curves, args.nq, seed=args.seed)
    vcurves, _, vIq = sur.generate(args.val, args.nq, seed=args.seed + 1000)
    t_gen = time.time() - t0
    print("generated %d + %d curves x %d q in %.1f s" % (args.curves, args.val, args.nq, t_gen), flush=True)

    raw = features(curves, lq)
    sur.set_normalisation(raw)
    X = torch.as_tensor(sur.standardise(raw), device="cuda")
    t = torch.as_tensor(target(curves, lq, Iq).ravel().astype(np.float32), device="cuda")
    t0 = time.time()
    hist = sur.train(X, t, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed,
                     log=lambda m: print(m, flush=True))
    torch.cuda.synchronize()
    t_train = time.time() - t0

    pred = sur.predict(vcurves, lq).reshape(args.val, args.nq).double().cpu().numpy()
    err = curve_errors(pred, vIq)
    summary = dict(curves=args.curves, nq=args.nq, epochs=args.epochs, batch=args.batch,
                   final_loss=hist[-1], train_s=t_train, gen_s=t_gen,
                   val_curves=args.val, val_max=float(err.max()), val_median=float(np.median(err)),
                   val_p90=float(np.percentile(err, 90)),
                   val_frac_below_1e3=float(np.mean(err <= 1e-3)),
                   val_frac_below_1e2=float(np.mean(err <= 1e-2)))
    print(json.dumps(summary), flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    sur.save(args.out)
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    return summary


if __name__ == "__main__":
    main()
