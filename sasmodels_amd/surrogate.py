"""
Learned surrogate for the polydisperse, orientation-averaged cylinder I(q).

The curve scaled by its nominal forward intensity, I / (1e-4 contrast^2
pi R^2 L), depends on q, R and L only through qR and qL (the lognormal
meshes scale with R and L).  An MLP (6 -> 4x256 SiLU -> 1) maps the features

    (log qR, log qL, radius_pd, length_pd, bL, bR),
    bL = log(1 + (qL/pi)^2)/2,   bR = 3/2 log(1 + (qR)^2/4)

(standardised) to a residual r, with

    I(q) = 1e-4 * contrast^2 * pi R^2 L * exp(r - bL - bR)   (scale 1, background 0)

bL and bR are smooth envelopes of the thin-rod and disc-cross-section
factors (pi/(qL) and (qR)^-3 at large argument), so r stays of order one
over the whole box and the BF16 layers carry only the correction.
Polydispersity is lognormal with 35 points over 3 sigma on radius and
length (none when the width is 0), and the orientation average is
Gauss-76 -- the plain kernel's definition.

Everything that computes runs in libsaship: the training data come from the
batched plain kernel (HipBatch), the forward and backward passes are the BF16
MFMA kernels of csrc/mlp.hip, inference is the fused kernel.  PyTorch holds
device memory and applies the Adam update to the flat FP32 parameter vector.

The plain kernel is also the fallback and the judge: inputs outside the
training box are computed by the plain kernel; every predicted curve is
checked at a few q points against the plain kernel and recomputed with it
when the relative error there exceeds the threshold.
"""

import ctypes as ct
import json
import math
import os

import numpy as np

from . import _lib, direct_model, generate
from .details import make_kernel_args
from .kernelhip import HipBatch, HipModel
from .modelinfo import make_model_info

#: training box: name -> (low, high)
BOX = {
    "radius": (10.0, 200.0),
    "length": (20.0, 2000.0),
    "radius_pd": (0.0, 0.3),
    "length_pd": (0.0, 0.3),
    "contrast": (0.1, 8.0),
}
LOG10Q = (-3.0, 0.0)
PD_N = 35
QUAD = 76
#: least fraction of held-out curves within a threshold for the weights to serve it
MIN_SERVED = 0.5


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the surrogate needs a GPU")
    return torch


def _stream(torch):
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


def curve_pars(c):
    """Model parameters (sld_solvent 0, sld = contrast) for a curve row
    (radius, length, radius_pd, length_pd, contrast)."""
    R, L, sR, sL, drho = [float(v) for v in c]
    return dict(radius=R, length=L, sld=drho, sld_solvent=0.0, scale=1.0, background=0.0,
                radius_pd=sR, radius_pd_n=PD_N, radius_pd_type="lognormal", radius_pd_nsigma=3,
                length_pd=sL, length_pd_n=PD_N, length_pd_type="lognormal", length_pd_nsigma=3)


def _raw(curves, log10q):
    """(features [n, nq, 6], baseline [n, nq]) in float64."""
    curves = np.asarray(curves, np.float64)
    lq = np.asarray(log10q, np.float64)[None, :]*np.log(10.0)
    lqr = lq + np.log(curves[:, 0])[:, None]
    lql = lq + np.log(curves[:, 1])[:, None]
    f = np.empty(lqr.shape + (6,))
    f[..., 0] = lqr
    f[..., 1] = lql
    f[..., 2] = curves[:, 2][:, None]
    f[..., 3] = curves[:, 3][:, None]
    f[..., 4] = 0.5*np.log1p((np.exp(lql)/np.pi)**2)
    f[..., 5] = 1.5*np.log1p(0.25*np.exp(lqr)**2)
    return f, -(f[..., 4] + f[..., 5])


def features(curves, log10q):
    """Raw (unstandardised) features for every (curve, q) row, float64 [n*nq, 6]."""
    f, _ = _raw(curves, log10q)
    return f.reshape(-1, 6)


def target(curves, log10q, Iq):
    """Network target r = log(I / (1e-4 contrast^2 pi R^2 L)) + bL + bR for I[n, nq]."""
    curves = np.asarray(curves, np.float64)
    v0 = np.pi*curves[:, 0]**2*curves[:, 1]
    _, base = _raw(curves, log10q)
    return np.log(np.asarray(Iq)/(1e-4*curves[:, 4, None]**2*v0[:, None])) - base


def iq_from_output(curves, log10q, r):
    """I(q) [n, nq] from network outputs r [n, nq]."""
    curves = np.asarray(curves, np.float64)
    v0 = np.pi*curves[:, 0]**2*curves[:, 1]
    _, base = _raw(curves, log10q)
    return 1e-4*curves[:, 4, None]**2*v0[:, None]*np.exp(np.asarray(r, np.float64) + base)


def sample_box(n, rng):
    """n curves uniform in the training box."""
    lo = np.array([BOX[k][0] for k in BOX])
    hi = np.array([BOX[k][1] for k in BOX])
    return lo + (hi - lo)*rng.random((n, 5))


def in_box(curves, log10q=None):
    curves = np.asarray(curves, np.float64)
    ok = np.ones(len(curves), bool)
    for j, k in enumerate(BOX):
        ok &= (curves[:, j] >= BOX[k][0]) & (curves[:, j] <= BOX[k][1])
    if log10q is not None:
        lq = np.asarray(log10q)
        if lq.min() < LOG10Q[0] - 1e-12 or lq.max() > LOG10Q[1] + 1e-12:
            ok[:] = False
    return ok


class PlainCylinder:
    """The plain HIP path for surrogate curves: batched, Gauss-76."""

    def __init__(self):
        info = make_model_info("cylinder")
        generate.set_integration_size(info, QUAD)
        self.info = info
        self.model = HipModel(info)

    def _kernel_like(self):
        return type("K", (), {"info": self.info, "dtype": np.dtype("d")})()

    def calls(self, curves):
        k = self._kernel_like()
        out = []
        for c in curves:
            mesh = direct_model.get_mesh(self.info, curve_pars(c), dim="1d")
            out.append(make_kernel_args(k, mesh)[:2])
        return out

    def batch(self, q_list, use_graph=False):
        return HipBatch(self.model, [[np.asarray(q, "d")] for q in q_list], use_graph=use_graph)

    def iq(self, curves, q_list, batch=None):
        """I(q) of every curve at its q points (list of arrays)."""
        own = batch is None
        b = self.batch(q_list) if own else batch
        out = b(self.calls(curves))
        if own:
            b.release()
        return out


class CylinderSurrogate:
    """Train, evaluate and gate the cylinder surrogate on device 0."""

    def __init__(self, seed=0):
        torch = _torch()
        self.torch = torch
        L = _lib.lib()
        n_params, n_pack, work = ct.c_int32(), ct.c_int32(), ct.c_int64()
        _lib.check(L.saship_mlp_sizes(ct.byref(n_params), ct.byref(n_pack), ct.byref(work), 1))
        self.n_params = n_params.value
        self.theta = torch.zeros(self.n_params, dtype=torch.float32, device="cuda")
        self.pack = torch.zeros(n_pack.value, dtype=torch.bfloat16, device="cuda")
        self.norm = torch.zeros(12, dtype=torch.float32, device="cuda")
        self._work = None
        self._work_batch = 0
        self.validation = None
        self.init_params(seed)
        self.plain = PlainCylinder()

    # ------------------------------------------------------------ parameters
    def init_params(self, seed=0):
        """He-normal weights, zero biases."""
        rng = np.random.default_rng(seed)
        D, H = 6, 256
        parts = [rng.normal(0, math.sqrt(2.0/D), D*H), np.zeros(H)]
        for _ in range(3):
            parts += [rng.normal(0, math.sqrt(2.0/H), H*H), np.zeros(H)]
        parts += [rng.normal(0, math.sqrt(1.0/H), H), np.zeros(1)]
        flat = np.concatenate(parts).astype(np.float32)
        assert flat.size == self.n_params
        self.theta.copy_(self.torch.from_numpy(flat))

    def set_normalisation(self, raw_features):
        mean = raw_features.mean(axis=0)
        std = raw_features.std(axis=0)
        std[std == 0] = 1.0
        self.norm.copy_(self.torch.from_numpy(np.concatenate([mean, std]).astype(np.float32)))

    def standardise(self, raw_features):
        nm = self.norm.cpu().numpy().astype(np.float64)
        return ((raw_features - nm[:6])/nm[6:]).astype(np.float32)

    def repack(self):
        _lib.check(_lib.lib().saship_mlp_pack(ct.c_void_p(self.theta.data_ptr()),
                                              ct.c_void_p(self.pack.data_ptr()), _stream(self.torch)))

    def _workspace(self, batch):
        if self._work is None or self._work_batch < batch:
            n_params, n_pack, work = ct.c_int32(), ct.c_int32(), ct.c_int64()
            _lib.check(_lib.lib().saship_mlp_sizes(ct.byref(n_params), ct.byref(n_pack), ct.byref(work), batch))
            self._work = self.torch.empty(work.value, dtype=self.torch.uint8, device="cuda")
            self._work_batch = batch
        return self._work

    # ------------------------------------------------------------ data
    def generate(self, ncurves=4096, nq=512, seed=1, chunk=512):
        """Curves uniform in the box at nq log-spaced q, computed by the plain kernel.
        Returns (curves [n,5], log10q [nq], Iq [n, nq])."""
        rng = np.random.default_rng(seed)
        curves = sample_box(ncurves, rng)
        log10q = np.linspace(LOG10Q[0], LOG10Q[1], nq)
        q = 10.0**log10q
        Iq = np.empty((ncurves, nq))
        b = self.plain.batch([q]*min(chunk, ncurves))
        for s in range(0, ncurves, chunk):
            part = curves[s:s + chunk]
            if len(part) != b.nsets:
                b.release()
                b = self.plain.batch([q]*len(part))
            Iq[s:s + len(part)] = np.array(b(self.plain.calls(part)))
        b.release()
        return curves, log10q, Iq

    # ------------------------------------------------------------ training
    def forward(self, X):
        """Network output for standardised features X (torch float32 [B, 6] on GPU)."""
        torch = self.torch
        B = X.shape[0]
        y = torch.empty(B, dtype=torch.float32, device="cuda")
        self.repack()
        _lib.check(_lib.lib().saship_mlp_forward(
            ct.c_void_p(self.theta.data_ptr()), ct.c_void_p(self.pack.data_ptr()), ct.c_void_p(X.data_ptr()),
            B, ct.c_void_p(self._workspace(B).data_ptr()), ct.c_void_p(y.data_ptr()), _stream(torch)))
        return y

    def loss_and_grad(self, X, t):
        """(mean squared error, gradient) for one batch, both on the GPU."""
        torch = self.torch
        B = X.shape[0]
        grad = torch.empty(self.n_params, dtype=torch.float32, device="cuda")
        parts = torch.empty((B + 255)//256, dtype=torch.float32, device="cuda")
        self.repack()
        _lib.check(_lib.lib().saship_mlp_train_step(
            ct.c_void_p(self.theta.data_ptr()), ct.c_void_p(self.pack.data_ptr()), ct.c_void_p(X.data_ptr()),
            ct.c_void_p(t.data_ptr()), B, ct.c_void_p(self._workspace(B).data_ptr()),
            ct.c_void_p(grad.data_ptr()), ct.c_void_p(parts.data_ptr()), _stream(torch)))
        return parts.sum()/B, grad

    def train(self, X, t, epochs=50, batch=8192, lr=2e-3, lr_end=1e-5, seed=2, log=None):
        """Adam on the flat parameters with a cosine learning-rate decay.
        X: standardised features [N, 6], t: targets [N] (torch, GPU)."""
        torch = self.torch
        param = torch.nn.Parameter(self.theta)
        opt = torch.optim.Adam([param], lr=lr)
        n = X.shape[0]
        steps_per_epoch = max(1, n//batch)
        total = epochs*steps_per_epoch
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        history = []
        step = 0
        for ep in range(epochs):
            perm = torch.randperm(n, device="cuda", generator=gen)
            acc = torch.zeros((), device="cuda")
            for s in range(steps_per_epoch):
                idx = perm[s*batch:(s + 1)*batch]
                loss, grad = self.loss_and_grad(X[idx].contiguous(), t[idx].contiguous())
                param.grad = grad
                for g in opt.param_groups:
                    g["lr"] = lr_end + 0.5*(lr - lr_end)*(1 + math.cos(math.pi*step/total))
                opt.step()
                acc += loss
                step += 1
            history.append(float(acc)/steps_per_epoch)
            if log:
                log("epoch %d loss %.3e" % (ep, history[-1]))
        self.theta = param.data
        return history

    # ------------------------------------------------------------ inference
    def predict(self, curves, log10q, out_iq=True, grid=0):
        """Surrogate output for curves [n,5] at log10q [nq]: I (out_iq) or y, torch [n*nq]."""
        torch = self.torch
        c = torch.as_tensor(np.asarray(curves, np.float32), device="cuda")
        lq = torch.as_tensor(np.asarray(log10q, np.float32), device="cuda")
        y = torch.empty(c.shape[0]*lq.shape[0], dtype=torch.float32, device="cuda")
        self.repack()
        self.infer_into(c, lq, y, out_iq, grid)
        return y

    def infer_into(self, curves_t, logq_t, out_t, out_iq=True, grid=0):
        """Fused inference on device tensors (no allocation, no host copies)."""
        _lib.check(_lib.lib().saship_mlp_infer(
            ct.c_void_p(self.theta.data_ptr()), ct.c_void_p(self.pack.data_ptr()),
            ct.c_void_p(curves_t.data_ptr()), int(curves_t.shape[0]), ct.c_void_p(logq_t.data_ptr()),
            int(logq_t.shape[0]), ct.c_void_p(self.norm.data_ptr()), 1 if out_iq else 0,
            ct.c_void_p(out_t.data_ptr()), int(grid), _stream(self.torch)))

    def serves(self, threshold):
        """
        Whether the loaded weights take requests at relative-error
        *threshold*: their validation record (the JSON written beside the
        weights by train_surrogate) must show at least MIN_SERVED of the
        held-out curves within it.  Below that most curves would be computed
        twice (inference, then the plain kernel), which is slower than the
        plain kernel alone.  Weights without a record (trained in this
        process) are served.
        """
        if not self.validation:
            return True
        key = "val_frac_below_1e3" if threshold < 1e-2 else "val_frac_below_1e2"
        return float(self.validation.get(key, 0.0)) >= MIN_SERVED

    def evaluate(self, curves, log10q, threshold=1e-3, judge_points=8, batch=None):
        """
        I(q) for every curve with the plain-kernel gate: (Iq [n, nq], used_fallback [n],
        judged_err [n]).  Curves outside the box, or whose relative error at
        `judge_points` log-spaced q exceeds `threshold`, come from the plain kernel;
        when the weights do not serve *threshold* (:meth:`serves`) every curve
        does, without inference (judged_err is then NaN).
        """
        curves = np.asarray(curves, np.float64)
        nq = len(log10q)
        q = 10.0**np.asarray(log10q, np.float64)
        if not self.serves(threshold):
            I = np.array(self.plain.iq(curves, [q]*len(curves))).reshape(len(curves), nq)
            return I, np.ones(len(curves), bool), np.full(len(curves), np.nan)
        I = self.predict(curves, log10q).reshape(len(curves), nq).double().cpu().numpy()
        jidx = np.unique(np.linspace(0, nq - 1, judge_points).round().astype(int))
        judge = np.array(self.plain.iq(curves, [q[jidx]]*len(curves)))
        err = np.max(np.abs(I[:, jidx] - judge)/np.abs(judge), axis=1)
        bad = (err > threshold) | ~in_box(curves, log10q)
        if bad.any():
            redo = self.plain.iq(curves[bad], [q]*int(bad.sum()))
            I[bad] = np.array(redo)
        return I, bad, err

    def save(self, path):
        np.savez(path, theta=self.theta.cpu().numpy(), norm=self.norm.cpu().numpy())

    def load(self, path):
        """Weights from *path* (.npz) and their validation record (the .json beside it, if any)."""
        d = np.load(path)
        self.theta = self.torch.as_tensor(d["theta"], device="cuda").contiguous()
        self.norm.copy_(self.torch.as_tensor(d["norm"]))
        record = os.path.splitext(path)[0] + ".json"
        self.validation = None
        if os.path.exists(record):
            with open(record) as f:
                self.validation = json.load(f)
