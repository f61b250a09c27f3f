"""
GPU tests of the surrogate: the BF16 MFMA forward and backward kernels
against the FP64 MLP oracle, inference against the training forward pass,
training on plain-kernel data, and the plain-kernel gate.
"""

import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sur():
    from sasmodels_amd.surrogate import CylinderSurrogate
    return CylinderSurrogate(seed=5)


def _batch(rng, B):
    return rng.normal(size=(B, 6)).astype(np.float32), rng.normal(size=B).astype(np.float32)


def test_forward_matches_oracle(sur):
    torch = sur.torch
    rng = np.random.default_rng(0)
    X, _ = _batch(rng, 300)
    y = sur.forward(torch.as_tensor(X, device="cuda")).cpu().numpy()
    theta = sur.theta.cpu().numpy()
    y_emul, _ = omlp.forward(theta, X, emulate_bf16=True)
    y_exact, _ = omlp.forward(theta, X, emulate_bf16=False)
    scale = np.abs(y_exact).max()
    assert np.max(np.abs(y - y_emul)) <= 2e-3*scale        # same roundings, FP32 sums
    assert np.max(np.abs(y - y_exact)) <= 5e-2*scale       # BF16 against exact FP64


def test_gradient_matches_oracle(sur):
    """Per parameter block: relative L2 error against the BF16-emulating
    FP64 oracle, and direction (cosine) against the exact FP64 gradient."""
    torch = sur.torch
    rng = np.random.default_rng(1)
    X, t = _batch(rng, 512)
    loss, grad = sur.loss_and_grad(torch.as_tensor(X, device="cuda"), torch.as_tensor(t, device="cuda"))
    grad = grad.cpu().numpy().astype(np.float64)
    theta = sur.theta.cpu().numpy()
    loss_e, g_emul = omlp.backward(theta, X, t, emulate_bf16=True)
    _, g_exact = omlp.backward(theta, X, t, emulate_bf16=False)
    assert abs(float(loss) - loss_e) <= 1e-3*loss_e
    off = 0
    for name, shape in omlp.SIZES:
        n = int(np.prod(shape))
        a, b, c = grad[off:off + n], g_emul[off:off + n], g_exact[off:off + n]
        off += n
        rel = np.linalg.norm(a - b)/max(np.linalg.norm(b), 1e-30)
        cos = np.dot(a, c)/max(np.linalg.norm(a)*np.linalg.norm(c), 1e-30)
        print(name, "rel %.2e cos %.6f" % (rel, cos))
        assert rel <= 2e-2, (name, rel)
        assert cos >= 0.999, (name, cos)


def test_inference_kernel_equals_training_forward(sur):
    torch = sur.torch
    rng = np.random.default_rng(2)
    from sasmodels_amd.surrogate import features, sample_box
    curves = sample_box(37, rng)
    lq = np.linspace(-3, 0, 61)
    raw = features(curves, lq)
    sur.set_normalisation(raw)
    X = torch.as_tensor(sur.standardise(raw), device="cuda")
    y_train = sur.forward(X).cpu().numpy()
    y_inf = sur.predict(curves, lq, out_iq=False).cpu().numpy()
    # The fused kernel builds the features on the device (FP32 log/exp) and the
    # training path on the host (FP64, then FP32): one-ulp differences in a
    # feature can flip a BF16 rounding of a layer input.  Almost every row
    # agrees to FP32 summation order; the rest differ by a few BF16 ulps.
    diff = np.abs(y_inf - y_train)
    scale = np.abs(y_train).max()
    assert np.mean(diff <= 1e-5*scale) >= 0.95
    assert diff.max() <= 1e-2*scale
    I = sur.predict(curves, lq, out_iq=True).cpu().numpy().reshape(37, 61)
    from sasmodels_amd.surrogate import iq_from_output
    np.testing.assert_allclose(I, iq_from_output(curves, lq, y_inf.reshape(37, 61)), rtol=5e-5)


def test_training_fits_plain_kernel_data_and_gate_falls_back(sur):
    """A short training on plain-kernel curves must cut the loss (of order
    one at initialisation: the target is the residual to the analytic
    envelope) by more than an order of magnitude; the gate must return
    plain-kernel curves where the surrogate is off or the input is outside
    the box."""
    torch = sur.torch
    from sasmodels_amd.surrogate import features, target
    curves, lq, Iq = sur.generate(ncurves=256, nq=128, seed=3)
    raw = features(curves, lq)
    sur.init_params(7)
    sur.set_normalisation(raw)
    X = torch.as_tensor(sur.standardise(raw), device="cuda")
    t = torch.as_tensor(target(curves, lq, Iq).ravel().astype(np.float32), device="cuda")
    hist = sur.train(X, t, epochs=30, batch=2048, lr=2e-3)
    print("loss", hist[0], hist[-1])
    assert hist[-1] < 0.05*hist[0] and hist[-1] < 0.05
    # gate: one curve outside the box, threshold 0 forces every curve to the plain kernel
    test = np.vstack([curves[:3], [[500.0, 100.0, 0.1, 0.1, 1.0]]])
    I, bad, err = sur.evaluate(test, lq, threshold=1e-3)
    assert bad[3]
    plain = np.array(sur.plain.iq(test, [10.0**lq]*4))
    np.testing.assert_allclose(I[bad], plain[bad], rtol=1e-12)
    I0, bad0, _ = sur.evaluate(test, lq, threshold=0.0)
    assert bad0.all()
    np.testing.assert_allclose(I0, plain, rtol=1e-12)


def test_trained_network_gains_over_the_envelope():
    """What the shipped network adds to the method without it: the analytic
    envelope alone (network output 0) against the trained network, both
    judged by the plain kernel on curves drawn from the box, in RMS of
    log I."""
    import os
    import torch
    from sasmodels_amd.surrogate import CylinderSurrogate, sample_box, target
    path = os.path.join(os.path.dirname(os.path.abspath(omlp.__file__)), "..", "sasmodels_amd", "data",
                        "cylinder_surrogate.npz")
    s = CylinderSurrogate(seed=0)
    s.load(path)
    s.repack()
    rng = np.random.default_rng(11)
    n, nq = 64, 128
    curves = sample_box(n, rng)
    lq = np.linspace(-3, 0, nq)
    plain = np.array(s.plain.iq(curves, [10.0**lq]*n))
    r_true = target(curves, lq, plain)
    out = torch.empty(n*nq, dtype=torch.float32, device="cuda")
    s.infer_into(torch.as_tensor(curves.astype(np.float32), device="cuda"),
                 torch.as_tensor(lq.astype(np.float32), device="cuda"), out, out_iq=False)
    r_net = out.view(n, nq).double().cpu().numpy()
    rms_envelope = float(np.sqrt(np.mean(r_true**2)))
    rms_network = float(np.sqrt(np.mean((r_net - r_true)**2)))
    print("RMS log I error: envelope alone %.4f, trained network %.4f" % (rms_envelope, rms_network))
    assert rms_network < 0.2*rms_envelope


def test_weights_that_miss_the_threshold_route_to_the_plain_kernel():
    """Weights whose validation record shows fewer than half the held-out
    curves within a threshold do not serve it: evaluate() returns the plain
    kernel's curves for every input without running inference."""
    import os
    from sasmodels_amd.surrogate import MIN_SERVED, CylinderSurrogate, sample_box
    path = os.path.join(os.path.dirname(os.path.abspath(omlp.__file__)), "..", "sasmodels_amd", "data",
                        "cylinder_surrogate.npz")
    s = CylinderSurrogate(seed=0)
    s.load(path)
    assert s.validation is not None
    served = s.validation["val_frac_below_1e3"] >= MIN_SERVED
    assert s.serves(1e-3) == served
    curves = sample_box(5, np.random.default_rng(4))
    lq = np.linspace(-3, 0, 40)
    I, bad, err = s.evaluate(curves, lq, threshold=1e-3)
    if not served:
        assert bad.all() and np.isnan(err).all()
        plain = np.array(s.plain.iq(curves, [10.0**lq]*5))
        np.testing.assert_allclose(I, plain, rtol=1e-12)
    s.validation = None          # weights without a record are served and judged
    assert s.serves(1e-3)
    I2, bad2, err2 = s.evaluate(curves, lq, threshold=1e-3)
    assert np.isfinite(err2).all()


def test_held_out_accuracy_matches_the_validation_record():
    """Inference I(q) of the shipped weights on fresh curves from the box,
    judged by the plain kernel: the per-curve maximum relative error (over
    q where I > 1e-6 of the curve's maximum) keeps to the validation record
    that decides which thresholds the weights serve (its median and its
    fraction of curves within 1e-2, with room for a different draw)."""
    import json
    import os
    import torch
    from sasmodels_amd.surrogate import CylinderSurrogate, sample_box
    data = os.path.join(os.path.dirname(os.path.abspath(omlp.__file__)), "..", "sasmodels_amd", "data")
    s = CylinderSurrogate(seed=0)
    s.load(os.path.join(data, "cylinder_surrogate.npz"))
    s.repack()
    with open(os.path.join(data, "cylinder_surrogate.json")) as f:
        rec = json.load(f)
    rng = np.random.default_rng(2027)
    n, nq = 128, 256
    curves = sample_box(n, rng)
    lq = np.linspace(-3, 0, nq)
    plain = np.array(s.plain.iq(curves, [10.0**lq]*n))
    out = torch.empty(n*nq, dtype=torch.float32, device="cuda")
    s.infer_into(torch.as_tensor(curves.astype(np.float32), device="cuda"),
                 torch.as_tensor(lq.astype(np.float32), device="cuda"), out, out_iq=True)
    I = out.view(n, nq).double().cpu().numpy()
    mask = plain > 1e-6*plain.max(axis=1, keepdims=True)
    err = np.where(mask, np.abs(I/plain - 1), 0.0).max(axis=1)
    median, within = float(np.median(err)), float(np.mean(err <= 1e-2))
    print("held-out: median per-curve max error %.4f (record %.4f), within 1e-2 %.2f (record %.2f)"
          % (median, rec["val_median"], within, rec["val_frac_below_1e2"]))
    assert np.isfinite(err).all()
    assert median <= 2.5*rec["val_median"]
    assert within >= 0.5*rec["val_frac_below_1e2"]
