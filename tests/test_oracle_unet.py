"""Pins the CPU network checker (oracle/conv_oracle.c): analytic gradients
against central finite differences; the forward pass reduced to closed forms by
parameter sets that make each layer a single tap (a shift with zero padding), an
identity through the average pools and nearest upsamplings, or a sum of the
concatenated channels (which fixes the [up, skip] order); Adam's closed-form
first step."""
import numpy as np
import pytest

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


# ---- the forward pass reduced to closed forms ----------------------------------------------

CENTER = 13  # tap (dz, dy, dx) = (0, 0, 0); tap t = (dz+1)*9 + (dy+1)*3 + (dx+1)


def tap(dz, dy, dx):
    return (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)


def theta_from(entries, biases=()):
    """parameters with W[layer][t][c][o] = value for (layer, t, c, o, value) in entries and
    bias[layer][o] = value for (layer, o, value) in biases, zero elsewhere (layout per layer:
    W[27][ci][co] then bias[co])"""
    theta = np.zeros(po.unet_nparams(2))
    offs, o = [], 0
    for ci, co in zip(CIN, COUT):
        offs.append(o)
        o += 27 * ci * co + co
    for l, t, c, oc, v in entries:
        theta[offs[l] + (t * CIN[l] + c) * COUT[l] + oc] = v
    for l, oc, v in biases:
        theta[offs[l] + 27 * CIN[l] * COUT[l] + oc] = v
    return theta


def positive_input(n, batch, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(batch, n, n, n, 2))
    x[..., 0] = rng.uniform(0.5, 1.5, size=(batch, n, n, n))  # ReLU passes channel 0
    return x


def block_mean(a, k):
    """mean over k^3 blocks, repeated back to full size (k x average pooling, then k x
    nearest upsampling)"""
    b, n = a.shape[0], a.shape[1]
    m = a.reshape(b, n // k, k, n // k, k, n // k, k).mean(axis=(2, 4, 6))
    return m.repeat(k, 1).repeat(k, 2).repeat(k, 3)


@pytest.mark.parametrize("d", [(0, 0, 0), (1, 0, -1), (-1, -1, -1), (0, 1, 1)])
def test_single_tap_is_a_shift_with_zero_padding(d):
    """one tap in the first layer and centre taps elsewhere, through the level-0 skip: the
    output is the input shifted by d, zero where v + d leaves the grid"""
    n, batch = 8, 2
    skip0 = 32  # concat [up (32 channels), skip (16)] before layer 7
    th = theta_from([(0, tap(*d), 0, 0, 1.0), (1, CENTER, 0, 0, 1.0),
                     (7, CENTER, skip0, 0, 1.0), (8, CENTER, 0, 0, 1.0)])
    x = positive_input(n, batch, 5)
    y = po.unet_forward(th, x.ravel(), batch, n, bf16=False).reshape(batch, n, n, n)
    ref = np.zeros((batch, n, n, n))
    dz, dy, dx = d
    src = x[..., 0]
    zs, ys, xs = (slice(max(0, -k), n - max(0, k)) for k in d)
    zd, yd, xd = (slice(max(0, k), n - max(0, -k)) for k in d)
    ref[:, zs, ys, xs] = src[:, zd, yd, xd]
    assert np.array_equal(y, ref)


def test_pooling_upsampling_and_concat_order():
    """identity layers along the coarse path and weights on both halves of each
    concatenation: y = up(up(pool(pool(x)))) + 2 up(pool(x)) + 3 x + bias, which fixes the
    pools as 2x2x2 averages, the upsamplings as nearest, and the concat order [up, skip]"""
    n, batch = 8, 2
    ent = [(l, CENTER, 0, 0, 1.0) for l in (0, 1, 2, 3, 4, 5)]
    ent += [(6, CENTER, 0, 0, 1.0), (6, CENTER, 64, 0, 2.0),   # [up(a5) 64 | skip a3 32]
            (7, CENTER, 0, 0, 1.0), (7, CENTER, 32, 0, 3.0),   # [up(a6) 32 | skip a1 16]
            (8, CENTER, 0, 0, 1.0)]
    th = theta_from(ent, biases=[(8, 0, 0.25)])
    x = positive_input(n, batch, 6)
    y = po.unet_forward(th, x.ravel(), batch, n, bf16=False).reshape(batch, n, n, n)
    x0 = x[..., 0]
    ref = block_mean(x0, 4) + 2 * block_mean(x0, 2) + 3 * x0 + 0.25
    assert np.allclose(y, ref, rtol=1e-14, atol=1e-14)


def test_relu_clips_negative_activations():
    """a negative first-layer weight: the ReLU after it passes nothing for a positive input,
    and the output is the output layer's bias alone"""
    n, batch = 4, 1
    th = theta_from([(0, CENTER, 0, 0, -1.0), (1, CENTER, 0, 0, 1.0), (7, CENTER, 32, 0, 1.0),
                     (8, CENTER, 0, 0, 1.0)], biases=[(8, 0, -0.5)])
    x = positive_input(n, batch, 7)
    y = po.unet_forward(th, x.ravel(), batch, n, bf16=False)
    assert np.array_equal(y, np.full(n ** 3, -0.5))
