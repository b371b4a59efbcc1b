"""The reference-style random boundary (FWI_RBC_GAUSSIAN) on the CPU.

The reference's boundary (src/modeling/modeling.cu:431-531,
src/admin/admin.cpp:115-206) is clock-seeded and cannot be reproduced, so
parity with it is pinned by what its algorithm fixes: Poisson-disk blobs at
least `ratio` apart in the outer half of the pad, radii in [0.2, 1] ratio,
amplitudes in [0.5, 1] varVp, one in three positive; Gaussian noise of
standard deviation f * varVp on the ramps; the interior untouched; values
clamped to [vmin - varVp, vmax + varVp].  The seeded statement is checked
bit for bit between the library's host sampler and the oracle's, and the
oracle's field against an independent numpy statement.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from fwi2d_amd import abi as A

f32 = np.float32

SHAPES = [  # nz, nx, nb, dh, ratio, amp, seed, shot, iteration
    (61, 181, 30, 10.0, 60.0, 500.0, 7, 3, 0),
    (301, 1001, 150, 10.0, 200.0, 1000.0, 20240, 17, 2),
    (81, 201, 41, 25.0, 200.0, 1000.0, 1, 0, 5),
    (40, 40, 8, 10.0, 35.0, 300.0, 3, 1, 1),
]


def oracle_blobs(nz, nx, nb, dh, ratio, amp, seed, shot, it):
    return ob.rb_blobs(nz + 2 * nb, nx + 2 * nb, nb, dh, ratio, amp, seed, (it << 32) | shot)


@pytest.mark.parametrize("shape", SHAPES)
def test_blob_sampler_matches_the_oracle(shape):
    nz, nx, nb, dh, ratio, amp, seed, shot, it = shape
    lib = A.rb_blobs(nz, nx, dh, nb, amp, ratio, seed, shot, it)
    ora = oracle_blobs(*shape)
    assert lib.shape == ora.shape and len(lib) > 0
    assert np.array_equal(lib.view(np.uint32), ora.view(np.uint32))


def test_no_blobs_without_radius_or_amplitude():
    assert len(A.rb_blobs(60, 80, 10.0, 20, 500.0, 0.0, 1, 0)) == 0
    assert len(A.rb_blobs(60, 80, 10.0, 20, 0.0, 200.0, 1, 0)) == 0
    assert len(oracle_blobs(60, 80, 20, 10.0, 200.0, 0.0, 1, 0, 0)) == 0


def test_blob_sampler_properties():
    nz, nx, nb, dh, ratio, amp = 201, 601, 100, 10.0, 150.0, 800.0
    x_max, z_max = f32(nx + 2 * nb - 1) * f32(dh), f32(nz + 2 * nb - 1) * f32(dh)
    h = 0.5 * nb * dh
    signs = []
    for shot in range(12):
        b = A.rb_blobs(nz, nx, dh, nb, amp, ratio, 20240, shot)
        x, z, c2, hh, a = b.T.astype(np.float64)
        # in the domain and in the outer half of the pad
        assert np.all((x >= 0) & (x < x_max) & (z >= 0) & (z < z_max))
        inner = (x > h) & (x < x_max - h) & (z > h) & (z < z_max - h)
        assert not inner.any()
        # Poisson disk: no two blobs closer than ratio
        d = np.sqrt((x[:, None] - x[None]) ** 2 + (z[:, None] - z[None]) ** 2)
        d[np.diag_indices(len(x))] = np.inf
        assert d.min() >= ratio * (1 - 1e-6)
        # radius in [0.2, 1] ratio, 16 r^2 and 0.5 / r^2 consistent
        r = np.sqrt(c2 / 16)
        assert r.min() >= 0.2 * ratio * (1 - 1e-6) and r.max() <= ratio * (1 + 1e-6)
        assert np.allclose(hh * c2, 8.0, rtol=1e-6)
        # amplitude in [0.5, 1] varVp
        assert np.abs(a).min() >= 0.5 * amp * (1 - 1e-6) and np.abs(a).max() <= amp
        signs.append(a > 0)
        # close to maximal: ring points of a 20 m lattice lie within 2 ratio of a blob
        gx, gz = np.meshgrid(np.arange(0, x_max, 20.0), np.arange(0, z_max, 20.0), indexing="ij")
        ring = ~((gx > h) & (gx < x_max - h) & (gz > h) & (gz < z_max - h))
        px, pz = gx[ring], gz[ring]
        near = np.zeros(len(px), bool)
        for k in range(len(x)):
            near |= (px - x[k]) ** 2 + (pz - z[k]) ** 2 < (2 * ratio) ** 2
        assert near.mean() > 0.995
        # deterministic, and another shot or iteration gives other blobs
        assert np.array_equal(b, A.rb_blobs(nz, nx, dh, nb, amp, ratio, 20240, shot))
    assert not np.array_equal(A.rb_blobs(nz, nx, dh, nb, amp, ratio, 20240, 0),
                              A.rb_blobs(nz, nx, dh, nb, amp, ratio, 20240, 0, 1))
    s = np.concatenate(signs)
    n = len(s)
    # positive with probability 1/3 (reference: rand() % 3 == 0), within 4 sigma
    assert abs(s.sum() - n / 3) < 4 * math.sqrt(n * (1 / 3) * (2 / 3))


def test_expneg_against_the_exponential():
    q = np.linspace(0, 8.5, 40001).astype(f32)
    got = np.array([ob.lib().or_rb_expneg(float(v)) for v in q], np.float64)
    ref = np.exp(-q.astype(np.float64))
    assert np.max(np.abs(got - ref) / ref) < 2.5e-7      # about 2 ulp
    assert ob.lib().or_rb_expneg(0.0) == 1.0


def test_normal_draws_are_irwin_hall():
    n = np.array([ob.lib().or_rb_normal(11, 7, c, k) for c in range(60000) for k in (0, 1)],
                 np.float64)
    assert abs(n.mean()) < 0.015
    assert abs(n.var() - 1.0) < 0.015
    kurt = np.mean((n - n.mean()) ** 4) / n.var() ** 2
    assert abs(kurt - 2.9) < 0.06            # Irwin-Hall(12): 3 - 6/(5*12)
    assert np.abs(n).max() <= 6.0
    # exact values: a multiple of 2^-17
    assert np.all((n * 131072) == np.round(n * 131072))


def fma32(a, b, c):
    # numpy has no fused multiply-add: the product of two float32 values is
    # exact in float64, the sum then rounds twice (float64, float32), which
    # can differ from one rounding by one unit in rare cases
    return f32(np.float64(a) * np.float64(b) + np.float64(c))


def expneg_py(q):
    n = int(f32(q * f32(1.44269504)))
    fn = f32(n)
    y = fma32(-fn, f32(0.693145752), q)
    y = fma32(-fn, f32(1.42860682e-6), y)
    p = f32(-2.75573192e-6)
    for cf in (2.48015873e-5, -1.98412698e-4, 1.38888889e-3, -8.33333333e-3, 4.16666667e-2,
               -1.66666667e-1, 0.5, -1.0, 1.0):
        p = fma32(p, y, f32(cf))
    return f32(np.ldexp(p, -n))


def field_statement(vpad, nb, dh, amp, ratio, vlo, vhi, seed, stream):
    """numpy statement of the FWI_RBC_GAUSSIAN field: the oracle's noise
    draws, the oracle's blob list, the bins and the clamp restated."""
    nxx, nzz = vpad.shape
    blobs = ob.rb_blobs(nzz, nxx, nb, dh, ratio, amp, seed, stream)
    x_max, z_max = f32(nxx - 1) * f32(dh), f32(nzz - 1) * f32(dh)
    binw = f32(4.0) * f32(ratio)
    nbx, nbz = int(x_max / binw) + 1, int(z_max / binw) + 1
    binof = lambda v, bw, nb_: min(int(f32(v) / bw), nb_ - 1)
    bins = {}
    for k, b in enumerate(blobs):
        bins.setdefault((binof(b[0], binw, nbx), binof(b[1], binw, nbz)), []).append(b)
    out = vpad.copy()
    for j in range(nxx):
        for i in range(nzz):
            xpad = j < nb or j >= nxx - nb
            zpad = i < nb or i >= nzz - nb
            if not (xpad or zpad):
                continue
            cell = i + j * nzz
            val = vpad[j, i]
            if xpad:
                dj = j if j < nb else nxx - 1 - j
                fs = f32(f32(1.0) - f32(f32(dj) / f32(nb))) * f32(amp)
                val = f32(val + f32(fs * f32(ob.lib().or_rb_normal(seed, stream, cell, 0))))
            if zpad:
                di = i if i < nb else nzz - 1 - i
                fs = f32(f32(1.0) - f32(f32(di) / f32(nb))) * f32(amp)
                val = f32(val + f32(fs * f32(ob.lib().or_rb_normal(seed, stream, cell, 1))))
            x, z = f32(j) * f32(dh), f32(i) * f32(dh)
            cbx, cbz = binof(x, binw, nbx), binof(z, binw, nbz)
            for bx in range(cbx - 1, cbx + 2):
                for bz in range(cbz - 1, cbz + 2):
                    for b in bins.get((bx, bz), []):
                        dx, dz = f32(x - b[0]), f32(z - b[1])
                        d2 = f32(f32(dx * dx) + f32(dz * dz))
                        if d2 < b[2]:
                            val = f32(val + f32(b[4] * expneg_py(f32(d2 * b[3]))))
            out[j, i] = min(max(val, f32(vlo)), f32(vhi))
    return out


def test_field_against_an_independent_statement():
    v = np.full((30, 24), 2000.0, np.float32)
    v[:, 12:] = 3000.0
    nb, dh, amp, ratio = 12, 10.0, 400.0, 40.0
    vpad = ob.expand(v, nb)
    vlo, vhi = f32(2000 - amp), f32(3000 + amp)
    for seed, stream in [(5, 0), (20240, (2 << 32) | 9)]:
        got = ob.rb_velocity_gauss(vpad, nb, dh, amp, ratio, vlo, vhi, seed, stream)
        exp = field_statement(vpad, nb, dh, amp, ratio, vlo, vhi, seed, stream)
        # fma32's double rounding agrees with one rounding on every operation
        # of these (fixed) inputs, so the statements agree bit for bit
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
        assert np.sum(np.abs(got - vpad) > 50) > 0.5 * got.size


def test_field_properties():
    nz, nx, nb, dh = 60, 90, 40, 10.0
    v = np.full((nx, nz), 2500.0, np.float32)
    vpad = ob.expand(v, nb)
    nxx, nzz = vpad.shape
    amp = 300.0
    vlo, vhi = f32(2500 - amp), f32(2500 + amp)
    # noise only (no blobs): (val - base) / (f amp) is a standard normal on
    # the x-pad columns (z inside), clamped at +-1 here
    noise = ob.rb_velocity_gauss(vpad, nb, dh, amp, 0.0, f32(0), f32(1e9), 9, 4)
    assert np.array_equal(noise[nb:-nb, nb:-nb], v)
    cols = np.arange(0, nb // 2)
    f = 1 - cols / nb
    u = (noise[cols, nb:-nb] - 2500.0) / (f[:, None] * amp)
    assert abs(u.mean()) < 0.05 and abs(u.std() - 1.0) < 0.05
    # with blobs: interior untouched, clamped to [vlo, vhi], and the cells
    # next to a blob centre carry its sign on average
    out = ob.rb_velocity_gauss(vpad, nb, dh, amp, 80.0, vlo, vhi, 9, 4)
    assert np.array_equal(out[nb:-nb, nb:-nb], v)
    assert out.min() >= vlo and out.max() <= vhi
    blobs = ob.rb_blobs(nzz, nxx, nb, dh, 80.0, amp, 9, 4)
    agree = []
    for x, z, c2, h, a in blobs:
        j, i = int(round(x / dh)), int(round(z / dh))
        pad = (j < nb or j >= nxx - nb or i < nb or i >= nzz - nb)
        if pad:
            agree.append(np.sign(out[j, i] - noise[j, i]) == np.sign(a) or out[j, i] in (vlo, vhi))
    assert np.mean(agree) > 0.9
    # deterministic in (seed, stream)
    assert np.array_equal(out, ob.rb_velocity_gauss(vpad, nb, dh, amp, 80.0, vlo, vhi, 9, 4))
    assert not np.array_equal(out, ob.rb_velocity_gauss(vpad, nb, dh, amp, 80.0, vlo, vhi, 9, 5))
