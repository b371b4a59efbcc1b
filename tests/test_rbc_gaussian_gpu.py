"""The reference-style random boundary (FWI_RBC_GAUSSIAN) on the GPU: the
field the engine's passes use equals the oracle's statement bit for bit
(host blob sampling, pinned upload and the generator kernel together), with
and without the no-allocation entry point, and whole passes run with it."""
import numpy as np
import pytest

import fwi2d_amd as F
import oracle_bind as ob
from fwi2d_amd import abi as A
from cases import overthrust, rel_l2

pytestmark = pytest.mark.gpu

f32 = np.float32


def bounds(v, amp, dh, dt):
    vlo = f32(v.min()) * f32(1.0) - f32(amp)
    vhi = f32(v.max()) * f32(1.0) + f32(amp)
    vcfl = (f32(0.999) * f32(0.5546) * f32(dh)) / f32(dt)
    return f32(vlo), f32(min(vhi, vcfl))


def engine(v, dh, dt, nb, amp, ratio, seed, max_batch=2, nt=401):
    nx, nz = v.shape
    eng = F.Engine(nz, nx, dh, nt, dt, 0, 0.0, nb, amp, 0.0, seed, max_batch,
                   rbc_kind=A.RBC_GAUSSIAN, rbc_ratio=ratio)
    w, tlag = F.ricker_wavelet(nt, dt, 20.0)
    eng.set_wavelet(w, tlag)
    xs = np.linspace(0.25, 0.75, 3) * (nx - 1) * dh
    xr = np.linspace(0, (nx - 1) * dh, 41)
    eng.set_geometry(f32(xs), np.full(3, 2 * dh, f32), f32(xr), np.full(41, dh, f32))
    eng.set_model(v)
    return eng


@pytest.mark.parametrize("ratio", [200.0, 0.0])
def test_gaussian_rb_velocity_matches_the_oracle(ratio):
    v = overthrust()
    dh, dt, nb, amp, seed = 25.0, 1e-3, 41, 1000.0, 20240
    print(f"overthrust 81x201, nb {nb}, varVp {amp}, ratio {ratio}", flush=True)
    eng = engine(v, dh, dt, nb, amp, ratio, seed)
    vpad = ob.expand(v, nb)
    vlo, vhi = bounds(v, amp, dh, dt)
    for shot, it in [(0, 0), (2, 0), (1, 7)]:
        ref = ob.rb_velocity_gauss(vpad, nb, dh, amp, ratio, vlo, vhi, seed, (it << 32) | shot)
        gpu = eng.rb_velocity(shot, it)
        assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32)), (shot, it)
        na = eng.rb_velocity(shot, it, noalloc=True)
        assert np.array_equal(na.view(np.uint32), ref.view(np.uint32)), (shot, it)
    a, b = eng.rb_velocity(0, 0), eng.rb_velocity(1, 0)
    assert not np.array_equal(a, b)
    assert np.array_equal(a[nb:-nb, nb:-nb], v)
    eng.close()


def test_gaussian_rb_batches_match_single_shots():
    """A migration with the Gaussian boundary: every slot of a batch samples
    and bins its own blobs, so batches of 2 and 3 and single shots give the
    same image bit for bit (the table is rebuilt and uploaded per batch)."""
    v = overthrust()
    dh, dt, nb, seed, nt = 25.0, 1e-3, 41, 7, 1001
    nx, nz = v.shape
    xs = np.linspace(1000, 4000, 4)
    xr = np.linspace(0, 5000, 51)
    w, tlag = F.ricker_wavelet(nt, dt, 30.0)

    def run(batch):
        eng = F.Engine(nz, nx, dh, nt, dt, 51, 0.004, nb, 1000.0, 0.0, seed, batch,
                       rbc_kind=A.RBC_GAUSSIAN, rbc_ratio=200.0)
        eng.set_wavelet(w, tlag)
        eng.set_geometry(f32(xs), np.full(4, 100.0, f32), f32(xr), np.zeros(51, f32))
        eng.set_model(v)
        obs = eng.modeling(F.ABC)
        eng.set_observed(obs)
        eng.reset()
        eng.migration()
        img, raw, ill = eng.get_image()
        eng.close()
        return raw, ill

    images = {}
    for batch in (1, 2, 3):
        print(f"migration, 4 shots in batches of {batch}, Gaussian boundary", flush=True)
        images[batch] = run(batch)
    raw, ill = images[1]
    assert np.isfinite(raw).all() and np.isfinite(ill).all() and ill.max() > 0
    for batch in (2, 3):
        assert np.array_equal(images[batch][0], raw) and np.array_equal(images[batch][1], ill)
