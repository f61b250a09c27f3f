"""
Pins the FP64 MLP oracle (oracle/mlp.py) that checks the surrogate's GPU
forward and backward kernels: its gradient against central finite
differences, and its BF16 rounding against torch's.
"""

import numpy as np
import pytest

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    theta = mlp.init(3).astype(np.float64)*0.5
    X = rng.normal(size=(7, mlp.D))
    t = rng.normal(size=7)
    _, g = mlp.backward(theta, X, t, emulate_bf16=False)
    layout = np.cumsum([0] + [int(np.prod(s)) for _, s in mlp.SIZES])
    picks = [layout[k] + j for k in range(len(mlp.SIZES)) for j in (0, 5, 17)
             if layout[k] + j < layout[k + 1]]
    h = 1e-6
    for i in picks:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (mlp.backward(tp, X, t, False)[0] - mlp.backward(tm, X, t, False)[0])/(2*h)
        assert abs(fd - g[i]) <= 1e-6*max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(size=5000), rng.normal(size=100)*1e-30,
                        np.array([1.0, 1.00390625, 1.0078125, -3.5e38])]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mlp.bf16(x), want)


def test_forward_shapes_and_exact_mode():
    rng = np.random.default_rng(2)
    theta = mlp.init(0)
    X = rng.normal(size=(5, mlp.D))
    y, cache = mlp.forward(theta, X, emulate_bf16=False)
    assert y.shape == (5,) and len(cache["Z"]) == 4
    y2, _ = mlp.forward(theta, X, emulate_bf16=True)
    assert np.max(np.abs(y2 - y)) < 0.05*np.max(np.abs(y)) + 1e-3
    assert mlp.NPARAM == 6*256 + 256 + 3*(256*256 + 256) + 256 + 1


def test_forward_matches_torch_float64():
    """The exact (FP64) forward against torch.nn.functional in float64 with
    the documented layout: W1[6][256], b1, W2..W4[256][256], b2..b4, w5[256], b5."""
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    rng = np.random.default_rng(4)
    theta = mlp.init(1).astype(np.float64) + rng.normal(0, 0.01, mlp.NPARAM)
    X = rng.normal(size=(9, mlp.D))
    y, _ = mlp.forward(theta, X, emulate_bf16=False)
    parts, off = [], 0
    for _, shape in mlp.SIZES:
        n = int(np.prod(shape))
        parts.append(torch.from_numpy(theta[off:off + n].reshape(shape)))
        off += n
    h = torch.from_numpy(X)
    for k in range(4):
        W, b = parts[2*k], parts[2*k + 1]
        h = F.silu(F.linear(h, W.T, b))
    ref = F.linear(h, parts[8].reshape(1, -1), parts[9]).reshape(-1).numpy()
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)
