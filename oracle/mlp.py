"""
FP64 numpy restatement of the surrogate MLP (6 -> 4x256 SiLU -> 1) with its
mean-squared-error gradient.

TEST INFRASTRUCTURE ONLY: the checker for the HIP forward and backward
kernels (sasmodels_amd/csrc/mlp.hip); never imported by the product.

``emulate_bf16=True`` rounds the operands of the three 256x256 layers to
BF16 (round to nearest even) at the points the GPU does -- layer inputs
and weights in the forward pass, the layer deltas and weights in the
backward pass -- and computes everything else in FP64.  With
``emulate_bf16=False`` the network is evaluated exactly in FP64.
The backward pass is pinned by central finite differences
(tests/test_oracle_mlp.py).
"""

import numpy as np

D, H = 6, 256
SIZES = [("W1", (D, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
         ("W4", (H, H)), ("b4", (H,)), ("w5", (H,)), ("b5", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unflatten(theta):
    theta = np.asarray(theta, np.float64)
    out, off = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = theta[off:off + n].reshape(shape)
        off += n
    return out


def silu(z):
    return z/(1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0/(1.0 + np.exp(-z))
    return s*(1.0 + z*(1.0 - s))


def init(seed=0):
    """He-style initialisation of the flat parameter vector (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SIZES:
        if name.startswith("W") or name == "w5":
            fan_in = shape[0]
            parts.append(rng.normal(0, np.sqrt(2.0/fan_in), shape).ravel())
        else:
            parts.append(np.zeros(shape).ravel())
    return np.concatenate(parts).astype(np.float32)


def forward(theta, X, emulate_bf16=True):
    """Network output y[B] and the cache for :func:`backward`."""
    p = unflatten(theta)
    rnd = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    X = np.asarray(X, np.float64)
    Z = [X @ p["W1"] + p["b1"]]
    A = [rnd(silu(Z[0]))]
    for l, name in enumerate(("W2", "W3", "W4")):
        Z.append(A[-1] @ rnd(p[name]) + p["b%d" % (l + 2)])
        A.append(rnd(silu(Z[-1])))
    y = silu(Z[3]) @ p["w5"] + p["b5"][0]
    return y, dict(X=X, Z=Z, A=A, p=p, rnd=rnd)


def backward(theta, X, target, emulate_bf16=True):
    """(loss, flat gradient) of mean((y - target)^2)."""
    y, c = forward(theta, X, emulate_bf16)
    p, Z, A, rnd = c["p"], c["Z"], c["A"], c["rnd"]
    B = len(y)
    r = y - np.asarray(target, np.float64)
    dy = 2.0*r/B
    g = {}
    g["w5"] = silu(Z[3]).T @ dy
    g["b5"] = np.array([dy.sum()])
    dZ = np.outer(dy, p["w5"])*dsilu(Z[3])
    for l in (2, 1, 0):                      # layers 4, 3, 2
        name = "W%d" % (l + 2)
        g[name] = A[l].T @ rnd(dZ)
        g["b%d" % (l + 2)] = dZ.sum(axis=0)
        dZ = (rnd(dZ) @ rnd(p[name]).T)*dsilu(Z[l])
    g["W1"] = c["X"].T @ dZ
    g["b1"] = dZ.sum(axis=0)
    flat = np.concatenate([g[name].ravel() for name, _ in SIZES])
    return float(np.mean(r*r)), flat
