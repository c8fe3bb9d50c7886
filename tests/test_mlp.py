"""The radiance-cache MLP oracle (oracle/mlp.py), pinned by
  - central finite differences of the FP64 loss (rel 1e-6);
  - the closed form of Adam's first step from zero moments (-lr sign(g));
  - exact bfloat16 rounding cases and the encoding's closed form;
and the C ABI's MLP entry points exist.  The GPU kernels are checked
against this oracle in test_gpu_mlp.py."""
import math

import numpy as np

from oracle import mlp


def _problem(n=8, seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    theta = mlp.init(seed)
    theta += rng.normal(scale=0.05, size=theta.size) * (np.arange(theta.size) >= 0)   # nonzero biases
    x = rng.uniform(0.0, 1.0, size=(n, mlp.N_IN))
    t = rng.uniform(0.0, 2.0, size=(n, mlp.N_OUT))
    return theta, x, t


def test_parameter_count():
    assert mlp.N_PARAMS == 28099
    assert mlp.FEAT == 240


def test_gradient_matches_central_differences():
    """With the loss weights held fixed the loss is quadratic in each
    parameter between ReLU kinks, so central differences are exact up to
    rounding; no pre-activation lies within 2e-3 of a kink."""
    theta, x, t = _problem(seed=5)
    h = mlp.encode(x)
    for W, b in mlp.unpack(theta)[:-1]:
        z = h @ W.T + b
        assert np.abs(z).min() > 2e-3
        h = np.maximum(z, 0.0)
    y, _ = mlp.forward(theta, x)
    w = mlp.loss_weights(y)

    def L(th):
        yy, _ = mlp.forward(th, x)
        return float(np.sum(w * (yy - t) ** 2))

    g = mlp.gradient(theta, x, t, weights=w)
    rng = np.random.Generator(np.random.PCG64(1))
    idx = np.concatenate([rng.choice(mlp.N_PARAMS, 300, replace=False),
                          np.arange(15360, 15370), np.arange(28090, 28099)])
    eps = 2e-4
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        fd = (L(tp) - L(tm)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), abs(g[i])) + 1e-9, (i, fd, g[i])


def test_adam_first_step_is_signed_learning_rate():
    theta, x, t = _problem()
    g = mlp.gradient(theta, x, t)
    th, m, v = theta.copy(), np.zeros_like(theta), np.zeros_like(theta)
    mlp.adam(th, m, v, g, 1)
    big = np.abs(g) > 1e-4
    np.testing.assert_allclose((th - theta)[big], -mlp.LR * np.sign(g[big]), rtol=1e-3)


def test_bf16_rounding():
    one = 1.0
    ulp = 2.0 ** -7                     # bfloat16 spacing at 1
    vals = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 4, -2.5, 3e38, 1e-40], dtype=np.float32)
    want = np.array([one, one, one + 2 * ulp, one, -2.5, np.float32(3e38), 0.0])
    got = mlp.bf16(vals)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == one + 2 * ulp and got[3] == 1.0 and got[4] == -2.5
    assert abs(got[5] - 3e38) / 3e38 < 2 ** -8
    assert abs(got[6]) < 1e-39


def test_encoding_closed_form():
    x = np.array([[0.25] + [0.0] * 9])
    f = mlp.encode(x)
    # input 0, band 0: sin(pi/4), cos(pi/4); band 1: sin(pi/2) = 1, cos = 0; band 2: sin(pi) = 0, cos = -1
    np.testing.assert_allclose(f[0, :6], [math.sqrt(0.5), math.sqrt(0.5), 1.0, 0.0, 0.0, -1.0], atol=1e-15)
    # input 1 (x = 0): sin 0 = 0, cos 0 = 1 in every band
    np.testing.assert_allclose(f[0, 24:48], np.tile([0.0, 1.0], 12), atol=0)


def test_abi_declares_mlp_entry_points(hip_lib):
    for name in ("tr_mlp_create", "tr_mlp_destroy", "tr_mlp_get_params", "tr_mlp_set_params", "tr_mlp_infer",
                 "tr_mlp_gradient", "tr_mlp_train_step"):
        assert hasattr(hip_lib, name)


def test_gpu_feature_recipe_matches_exact_encoding():
    """The angle-doubling features the GPU computes stay within 2e-6 of the
    exact sin/cos (three binary32 doublings from an exact base), far below
    the BF16 spacing of the features; a wrong doubling formula or band
    order would be off by O(1)."""
    x = np.random.default_rng(3).uniform(0.0, 1.0, size=(4096, mlp.N_IN))
    exact, gpu = mlp.encode(x.astype(np.float32)), mlp.encode_gpu(x)
    assert np.abs(gpu - exact).max() < 2e-6
    # the first band of each group of four is sin/cos itself
    f = gpu.reshape(-1, mlp.N_IN, mlp.BANDS, 2)
    e = exact.reshape(-1, mlp.N_IN, mlp.BANDS, 2)
    assert np.abs(f[:, :, ::4] - e[:, :, ::4]).max() < 2e-7
