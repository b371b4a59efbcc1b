"""Pins the FP64 network oracle: brute-force convolution, finite-difference
gradients of the whole network, and the bf16 rounding rule."""
import numpy as np

import nn_oracle as no


def brute_conv(x, W, b):
    B, C, H, Wd = x.shape
    y = np.zeros((B, W.shape[0], H, Wd))
    for bb in range(B):
        for o in range(W.shape[0]):
            for h in range(H):
                for w in range(Wd):
                    s = b[o]
                    for ky in range(3):
                        for kx in range(3):
                            hh, ww = h + ky - 1, w + kx - 1
                            if 0 <= hh < H and 0 <= ww < Wd:
                                s += np.dot(W[o, ky, kx, :], x[bb, :, hh, ww])
                    y[bb, o, h, w] = s
    return y


def test_conv_matches_brute_force_and_its_adjoint():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 5, 6))
    W = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    y = no.conv(x, W, b)
    assert np.abs(y - brute_conv(x, W, b)).max() < 1e-12
    # <conv(x), dy> = <x, conv^T(dy)> and dW, db by linearity
    dy = rng.standard_normal(y.shape)
    dx, dW, db = no.conv_backward(x, W, dy)
    lhs = np.sum((y - b[None, :, None, None]) * dy)
    assert abs(lhs - np.sum(x * dx)) < 1e-9 * abs(lhs)
    assert abs(lhs - np.sum(W * dW)) < 1e-9 * abs(lhs)
    assert np.allclose(db, dy.sum(axis=(0, 2, 3)))


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    n = sum(cout * 9 * cin + cout for cin, cout in no.ARCH)
    return rng.standard_normal(n) * scale / 4


def test_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = random_params(2)
    x = rng.standard_normal((2, 3, 8, 8))
    t = rng.standard_normal((2, 8, 8))
    loss, g, _ = no.loss_and_grad(p, x, t)
    idx = rng.choice(p.size, 40, replace=False)
    # include the first / last layer and a bias
    idx = np.concatenate([idx, [0, p.size - 1, p.size - 2, 32 * 9 * 3 + 5]])
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (no.loss_and_grad(p + e, x, t)[0] - no.loss_and_grad(p - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g).max()), (k, fd, g[k])


def test_bf16_rounding_rule():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, 1e-40], np.float64)
    r = no.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                     # halfway below an even mantissa -> down
    assert r[2] == 1.0 + 2 ** -7           # above halfway -> up
    assert r[3] == -2.5
    bits = np.asarray(r, np.float32).view(np.uint32)
    assert np.all(bits & 0xFFFF == 0)
