"""Pins the CPU network checker (oracle/conv_oracle.c): analytic gradients
against central finite differences, equivariance checks of the convolution,
and Adam's closed-form first step."""
import numpy as np

from oracle import pyoracle as po


CIN = [2, 16, 16, 32, 32, 64, 96, 48, 16]
COUT = [16, 16, 32, 32, 64, 64, 32, 16, 1]


def setup(n=4, batch=2, seed=0):
    """He-scaled weights (activations of order one), small random biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in zip(CIN, COUT):
        parts.append(rng.normal(scale=np.sqrt(2.0 / (27 * ci)), size=27 * ci * co))
        parts.append(rng.normal(scale=0.1, size=co))
    theta = np.concatenate(parts)
    assert theta.size == po.unet_nparams(2)
    x = rng.normal(size=batch * n ** 3 * 2)
    t = rng.normal(size=batch * n ** 3)
    return theta, x, t


def test_nparams():
    # 27*(2*16 + 16*16 + 16*32 + 32*32 + 32*64 + 64*64 + 96*32 + 48*16 + 16*1) + biases
    w = 27 * (32 + 256 + 512 + 1024 + 2048 + 4096 + 3072 + 768 + 16)
    b = 16 + 16 + 32 + 32 + 64 + 64 + 32 + 16 + 1
    assert po.unet_nparams(2) == w + b


def test_gradient_finite_differences():
    n, batch = 4, 2
    theta, x, t = setup(n, batch)
    loss, g = po.unet_loss_grad(theta, x, t, batch, n, bf16=False)
    rng = np.random.default_rng(1)
    idx = rng.choice(theta.size, 40, replace=False)
    # include the first and last layers' parameters
    idx = np.concatenate([idx, [0, 5, theta.size - 1, theta.size - 2]])
    eps = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        lp, _ = po.unet_loss_grad(tp, x, t, batch, n, bf16=False)
        lm, _ = po.unet_loss_grad(tm, x, t, batch, n, bf16=False)
        fd = (lp - lm) / (2 * eps)
        # central differences: O(eps^2) truncation, O(|loss| 1e-16 / eps) roundoff
        assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(g).max()), (i, fd, g[i])


def test_output_layer_is_affine():
    """The last convolution has no activation: doubling its weights and bias
    doubles the network's output exactly."""
    n, batch = 8, 1
    theta, x, _ = setup(n, batch, 3)
    y1 = po.unet_forward(theta, x, batch, n, bf16=False)
    th2 = theta.copy()
    th2[-(27 * 16 + 1):] *= 2.0  # last conv weights and bias: output doubles
    y2 = po.unet_forward(th2, x, batch, n, bf16=False)
    assert np.allclose(y2, 2 * y1, rtol=1e-12, atol=1e-12)


def test_bf16_rounding_is_small():
    n, batch = 4, 1
    theta, x, _ = setup(n, batch, 4)
    y = po.unet_forward(theta, x, batch, n, bf16=False)
    yb = po.unet_forward(theta, x, batch, n, bf16=True)
    assert np.abs(yb - y).max() <= 0.05 * np.abs(y).max()
    assert np.abs(yb - y).max() > 0


def test_adam_first_step():
    """After one bias-corrected step, every parameter moves by lr*sign(g)
    (up to eps)."""
    rng = np.random.default_rng(2)
    th = rng.normal(size=100)
    g = rng.normal(size=100)
    t1, m, v = po.adam(th, g, np.zeros(100), np.zeros(100), 1, lr=1e-3)
    assert np.allclose(th - t1, 1e-3 * np.sign(g), rtol=1e-6, atol=1e-9)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)
