"""FP64 numpy statement of the learned-preconditioner network (test oracle).

Same architecture as include/fwi2d_nn.h: 8 3x3 same-padding convolutions,
two 2x2 average pools, two nearest 2x upsamplings with additive skips,
ReLU everywhere except the last layer, mean-squared-error loss.  Arrays are
[batch][channel][H][W].  With bf16=True the forward and backward values are
rounded to bfloat16 (round-to-nearest-even) at the points where the GPU
stores bf16 tensors, so the GPU result can be compared tightly.

Test infrastructure only (tests import it as the checker).
"""
import numpy as np

ARCH = [(3, 32), (32, 32), (32, 64), (64, 64), (64, 64), (64, 32), (32, 32), (32, 1)]


def bf16(a):
    """Round to bfloat16 (nearest even) and return float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unflatten(params):
    """Flat params (W[cout][3][3][cin] then b[cout] per layer) -> list of (W, b)."""
    out, o = [], 0
    for cin, cout in ARCH:
        n = cout * 9 * cin
        W = np.asarray(params[o:o + n], np.float64).reshape(cout, 3, 3, cin)
        o += n
        b = np.asarray(params[o:o + cout], np.float64)
        o += cout
        out.append((W, b))
    assert o == len(params)
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def conv(x, W, b):
    """y[b,co,h,w] = sum W[co,ky,kx,ci] x[b,ci,h+ky-1,w+kx-1] + bias[co]."""
    B, C, H, Wd = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    y = np.zeros((B, W.shape[0], H, Wd))
    for ky in range(3):
        for kx in range(3):
            y += np.einsum("oc,bchw->bohw", W[:, ky, kx, :], xp[:, :, ky:ky + H, kx:kx + Wd])
    return y + b[None, :, None, None]


def conv_backward(x, W, dy):
    """Returns (dx, dW, db) of conv for upstream gradient dy."""
    B, C, H, Wd = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    dW = np.zeros_like(W)
    dxp = np.zeros_like(xp)
    for ky in range(3):
        for kx in range(3):
            dW[:, ky, kx, :] = np.einsum("bohw,bchw->oc", dy, xp[:, :, ky:ky + H, kx:kx + Wd])
            dxp[:, :, ky:ky + H, kx:kx + Wd] += np.einsum("oc,bohw->bchw", W[:, ky, kx, :], dy)
    return dxp[:, :, 1:-1, 1:-1], dW, dy.sum(axis=(0, 2, 3))


def pool(x):
    B, C, H, W = x.shape
    return x.reshape(B, C, H // 2, 2, W // 2, 2).mean(axis=(3, 5))


def pool_backward(d):
    return np.repeat(np.repeat(d, 2, axis=2), 2, axis=3) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=2), 2, axis=3)


def up_backward(d):
    B, C, H, W = d.shape
    return d.reshape(B, C, H // 2, 2, W // 2, 2).sum(axis=(3, 5))


def forward(params, x, rnd=False):
    """Returns (y [B,H,W], cache)."""
    R = bf16 if rnd else (lambda a: a)
    L = unflatten(params)
    if rnd:
        L = [(bf16(W), b) for W, b in L]
    relu = lambda a: np.maximum(a, 0.0)
    c = {}
    c["x0"] = R(x)
    c["a1"] = R(relu(conv(c["x0"], *L[0])))
    c["a2"] = R(relu(conv(c["a1"], *L[1])))
    c["p2"] = R(pool(c["a2"]))
    c["a3"] = R(relu(conv(c["p2"], *L[2])))
    c["a4"] = R(relu(conv(c["a3"], *L[3])))
    c["p4"] = R(pool(c["a4"]))
    c["a5"] = R(relu(conv(c["p4"], *L[4])))
    c["s5"] = R(up(c["a5"]) + c["a3"])
    c["a6"] = R(relu(conv(c["s5"], *L[5])))
    c["s6"] = R(up(c["a6"]) + c["a1"])
    c["a7"] = R(relu(conv(c["s6"], *L[6])))
    y = conv(c["a7"], *L[7])[:, 0]
    c["L"] = L
    return y, c


def loss_and_grad(params, x, target, rnd=False):
    """MSE loss and d loss / d params (flat, same layout as params)."""
    R = bf16 if rnd else (lambda a: a)
    y, c = forward(params, x, rnd)
    L = c["L"]
    n = y.size
    loss = float(np.mean((y - target) ** 2))
    mask = lambda a: (a > 0).astype(np.float64)
    dy = R((2.0 / n) * (y - target))[:, None]
    g = [None] * 8
    dx, dW, db = conv_backward(c["a7"], L[7][0], dy)
    g[7] = (dW, db)
    dz7 = R(dx * mask(c["a7"]))
    dx, dW, db = conv_backward(c["s6"], L[6][0], dz7)
    g[6] = (dW, db)
    ds6 = R(dx)
    dz6 = R(up_backward(ds6) * mask(c["a6"]))
    dx, dW, db = conv_backward(c["s5"], L[5][0], dz6)
    g[5] = (dW, db)
    ds5 = R(dx)
    dz5 = R(up_backward(ds5) * mask(c["a5"]))
    dx, dW, db = conv_backward(c["p4"], L[4][0], dz5)
    g[4] = (dW, db)
    dp4 = R(dx)
    dz4 = R(pool_backward(dp4) * mask(c["a4"]))
    dx, dW, db = conv_backward(c["a3"], L[3][0], dz4)
    g[3] = (dW, db)
    dz3 = R((dx + ds5) * mask(c["a3"]))
    dx, dW, db = conv_backward(c["p2"], L[2][0], dz3)
    g[2] = (dW, db)
    dp2 = R(dx)
    dz2 = R(pool_backward(dp2) * mask(c["a2"]))
    dx, dW, db = conv_backward(c["a1"], L[1][0], dz2)
    g[1] = (dW, db)
    dz1 = R((dx + ds6) * mask(c["a1"]))
    _, dW, db = conv_backward(c["x0"], L[0][0], dz1)
    g[0] = (dW, db)
    return loss, flatten(g), y
