"""GPU tests of the learned-preconditioner network through the C ABI:
the MFMA lane layout, forward and backward against the FP64 oracle (with the
GPU's bf16 rounding points emulated), the Adam update in closed form, and a
training run that reduces the loss."""
import numpy as np
import pytest

import fwi2d_amd as F
import nn_oracle as no
from cases import rel_l2

pytestmark = pytest.mark.gpu


def test_mfma_lane_layout():
    rng = np.random.default_rng(0)
    a = no.bf16(rng.standard_normal((16, 32))).astype(np.float32)
    b = no.bf16(rng.standard_normal((32, 16))).astype(np.float32)
    d = F.mfma_probe(a, b)
    ref = a.astype(np.float64) @ b.astype(np.float64)
    assert np.abs(d - ref).max() <= 1e-5 * np.abs(ref).max()


def small_problem(seed=0, nb=2, h=16, w=24):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((nb, 3, h, w)).astype(np.float32)
    t = (0.1 * rng.standard_normal((nb, h, w))).astype(np.float32)
    return x, t


def test_forward_matches_oracle():
    x, _ = small_problem()
    net = F.Net(2, 16, 24, seed=7)
    p = net.get_params()
    y = net.forward(x)
    yo, _ = no.forward(p.astype(np.float64), x.astype(np.float64), rnd=True)
    err = rel_l2(y, yo)
    print(f"forward relL2 (bf16 emulated) {err:.3e}")
    assert err < 1e-6
    yo64, _ = no.forward(p.astype(np.float64), x.astype(np.float64), rnd=False)
    # bf16 activations against an FP64 network: a few units of bf16's 2^-8
    print(f"forward relL2 (exact FP64) {rel_l2(y, yo64):.3e}")
    assert rel_l2(y, yo64) < 2e-2


def test_gradient_matches_oracle():
    x, t = small_problem(1)
    net = F.Net(2, 16, 24, seed=3)
    p = net.get_params()
    loss, g = net.gradient(x, t)
    lo, go, _ = no.loss_and_grad(p.astype(np.float64), x.astype(np.float64),
                                 t.astype(np.float64), rnd=True)
    print(f"loss {loss:.8e} oracle {lo:.8e}")
    assert abs(loss - lo) <= 1e-5 * lo
    # per layer relative error of the weight and bias gradients
    o = 0
    for k, (cin, cout) in enumerate(no.ARCH):
        n = cout * 9 * cin + cout
        e = rel_l2(g[o:o + n], go[o:o + n])
        print(f"layer {k + 1} gradient relL2 {e:.3e}")
        assert e < 1e-3, k
        o += n


def test_gradient_matches_finite_differences_per_layer():
    """The GPU gradient against central differences of the GPU loss, one
    direction per layer: u = the layer's gradient, normalised; the difference
    (L(p + eps u) - L(p - eps u)) / 2 eps must equal <g, u> within 5 %.  The
    loss of a bf16 network moves in bf16 steps, so eps is 0.3 % of the layer's
    weight norm, where the difference stands above that rounding (the
    bf16-emulated oracle gives ratios within 2.8 % of 1 for these parameters
    and inputs)."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 16, 24)).astype(np.float32)
    t = (0.1 * rng.standard_normal((2, 16, 24))).astype(np.float32)
    parts = []
    for cin, cout in no.ARCH:                     # He-uniform weights, zero biases
        lim = np.sqrt(6.0 / (9 * cin))
        parts += [rng.uniform(-lim, lim, cout * 9 * cin), np.zeros(cout)]
    p = np.concatenate(parts).astype(np.float32).astype(np.float64)
    net = F.Net(2, 16, 24, seed=0)
    net.set_params(p.astype(np.float32))
    _, g = net.gradient(x, t)
    g = g.astype(np.float64)
    o = 0
    for k, (cin, cout) in enumerate(no.ARCH):
        n = cout * 9 * cin + cout
        u = np.zeros_like(p)
        u[o:o + n] = g[o:o + n]
        u /= np.linalg.norm(u)
        eps = 3e-3 * np.linalg.norm(p[o:o + n - cout])
        net.set_params((p + eps * u).astype(np.float32))
        lp = net.gradient(x, t)[0]
        net.set_params((p - eps * u).astype(np.float32))
        lm = net.gradient(x, t)[0]
        fd = (lp - lm) / (2 * eps)
        gd = float(g @ u)
        print(f"layer {k + 1}: finite difference {fd:.5e}, <g, u> {gd:.5e}, ratio {fd / gd:.4f}")
        assert abs(fd / gd - 1.0) <= 0.05, k
        o += n
    net.close()


def test_adam_step_closed_form():
    x, t = small_problem(2)
    net = F.Net(2, 16, 24, seed=5)
    p0 = net.get_params()
    _, g = net.gradient(x, t)
    lr = 1e-3
    net.train_step(x, t, lr)
    p1 = net.get_params()
    # first Adam step from zero moments: w -= lr * g / (|g| + eps) (bias-corrected)
    expect = p0 - lr * g / (np.abs(g) + 1e-8)
    assert np.abs(p1 - expect).max() <= 1e-6 + 1e-4 * lr


def test_set_get_roundtrip_and_training_reduces_loss():
    x, t = small_problem(3, nb=4)
    net = F.Net(4, 16, 24, seed=11)
    p = net.get_params()
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    losses = [net.train_step(x, t, 3e-3) for _ in range(60)]
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.5 * losses[0]


def test_optimizer_state_closed_form_and_resume():
    """After one Adam step from zero moments m = (1 - b1) g, v = (1 - b2) g^2
    (b1 = 0.9, b2 = 0.999 as float32) and the step count is 1; a fresh network given the parameters and Adam
    state continues exactly as the original one does."""
    x, t = small_problem(4, nb=2)
    a = F.Net(2, 16, 24, seed=7)
    _, g = a.gradient(x, t)
    a.train_step(x, t, 2e-3)
    m, v, steps = a.get_optimizer()
    assert steps == 1
    # the kernel's betas are the float32 constants 0.9f and 0.999f
    b1, b2 = float(np.float32(0.9)), float(np.float32(0.999))
    assert np.allclose(m, (1.0 - b1) * g.astype(np.float64), rtol=1e-6, atol=0)
    assert np.allclose(v, (1.0 - b2) * g.astype(np.float64) ** 2, rtol=1e-6, atol=0)
    params = a.get_params()
    b = F.Net(2, 16, 24, seed=99)
    b.set_params(params)
    b.set_optimizer(m, v, steps)
    for _ in range(3):
        la, lb = a.train_step(x, t, 2e-3), b.train_step(x, t, 2e-3)
        assert la == lb
    assert np.array_equal(a.get_params(), b.get_params())
    ma, va, sa = a.get_optimizer()
    mb, vb, sb = b.get_optimizer()
    assert sa == sb == 4 and np.array_equal(ma, mb) and np.array_equal(va, vb)
    # invalid state is refused before it reaches the device
    with pytest.raises(F.FwiError):
        b.set_optimizer(m, -np.abs(v) - 1.0, steps)
    with pytest.raises(F.FwiError):
        b.set_optimizer(m, v, -1)
    with pytest.raises(ValueError):
        b.set_optimizer(m[:-1], v, steps)
    mc, vc, sc = b.get_optimizer()            # a refused call changes nothing
    assert sc == sb and np.array_equal(mc, mb) and np.array_equal(vc, vb)
    a.close()
    b.close()
