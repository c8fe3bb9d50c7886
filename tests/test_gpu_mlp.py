"""GPU radiance-cache MLP against the FP64 oracle (oracle/mlp.py).

The GPU multiplies BF16 operands with FP32 accumulation; the oracle's bf16
mode rounds the same operands in FP64, with features made by the GPU's
FP32 recipe (oracle.mlp.encode_gpu).  They differ by the accumulation
order and by rare one-ulp BF16 rounding flips (the GPU's first band of a
group comes from the shared FP32 sin/cos, the oracle's from FP64), so
outputs are compared to 1e-2 of their scale and gradients to 5e-4 in
norm, per layer.
"""
import numpy as np
import pytest

from oracle import mlp
from tempestrenderer_amd.scene import RadianceCacheMLP

pytestmark = pytest.mark.gpu


def _setup(n, seed=5):
    rng = np.random.Generator(np.random.PCG64(seed))
    theta = mlp.init(seed) + rng.normal(scale=0.05, size=mlp.N_PARAMS)
    theta32 = theta.astype(np.float32)
    x = rng.uniform(0.0, 1.0, size=(n, mlp.N_IN)).astype(np.float32)
    t = rng.uniform(0.0, 2.0, size=(n, mlp.N_OUT)).astype(np.float32)
    net = RadianceCacheMLP(seed=1)
    net.params = theta32
    return net, theta32.astype(np.float64), x, t


def _layer_slices():
    out, o = [], 0
    for fi, fo in mlp.LAYERS:
        out.append(slice(o, o + fi * fo + fo))
        o += fi * fo + fo
    return out


def test_params_round_trip():
    print("case: mlp params round trip", flush=True)
    net, theta, _, _ = _setup(8)
    assert np.array_equal(net.params, theta.astype(np.float32))


def test_inference_matches_oracle():
    print("case: mlp inference 4096", flush=True)
    net, theta, x, _ = _setup(4096)
    y = net.infer(x).astype(np.float64)
    ref, _ = mlp.forward(theta, x.astype(np.float64), bf16_mode=True)
    scale = np.abs(ref).max()
    err = np.abs(y - ref)
    assert err.max() <= 1e-2 * scale, (err.max(), scale)
    assert err.mean() <= 1e-3 * scale, (err.mean(), scale)
    ref64, _ = mlp.forward(theta, x.astype(np.float64))
    print("bf16 vs fp64 network: max %.3e mean %.3e of scale %.3f" % (
        np.abs(y - ref64).max(), np.abs(y - ref64).mean(), scale))


def test_gradient_matches_oracle():
    print("case: mlp gradient 2048", flush=True)
    net, theta, x, t = _setup(2048)
    valid = (np.arange(2048) % 7 != 3).astype(np.float32)
    g, loss = net.gradient(x, t, valid)
    ref = mlp.gradient(theta, x.astype(np.float64), t.astype(np.float64), valid=valid, bf16_mode=True,
                       loss_scale=128.0)
    ref_loss = mlp.loss(theta, x.astype(np.float64), t.astype(np.float64), valid=valid, bf16_mode=True)
    assert abs(loss - ref_loss) <= 1e-2 * ref_loss, (loss, ref_loss)
    for i, sl in enumerate(_layer_slices()):
        rel = np.linalg.norm(g[sl] - ref[sl]) / np.linalg.norm(ref[sl])
        print("layer %d gradient rel err %.2e" % (i, rel))
        assert rel <= 5e-4, (i, rel)


def _smooth_targets(x):
    return np.stack([np.sin(2 * np.pi * x[:, 0]) + 1.2, x[:, 1] * x[:, 2] + 0.1,
                     np.cos(np.pi * x[:, 3]) * 0.5 + 0.6], axis=1).astype(np.float32)


def test_adam_training_step_and_convergence():
    print("case: mlp training 300 steps", flush=True)
    net, theta, x, _ = _setup(4096)
    t = _smooth_targets(x)
    ref_g = mlp.gradient(theta, x.astype(np.float64), t.astype(np.float64), bf16_mode=True, loss_scale=128.0)
    before = net.params.astype(np.float64)
    loss0 = net.train_step(x, t)
    step = net.params.astype(np.float64) - before
    big = np.abs(ref_g) > 1e-3 * np.abs(ref_g).max()
    agree = np.mean(np.sign(step[big]) == -np.sign(ref_g[big]))
    assert agree > 0.99, agree
    np.testing.assert_allclose(np.abs(step[big]), mlp.LR, rtol=1e-3)
    losses = [loss0]
    for _ in range(299):
        losses.append(net.train_step(x, t))
    print("loss %.4f -> %.4f" % (losses[0], losses[-1]))
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.1 * losses[0]
    y = net.infer(x)
    assert np.abs(y - t).mean() < 0.1
