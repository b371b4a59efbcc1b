"""GPU parity: the HIP path through the C ABI against the FP64 CPU oracle.

Tolerances (relative L2 against the FP64 oracle on identical inputs):
  seismograms 1e-5, images / gradients / illumination 1e-4,
  random-boundary velocity: bit-exact, launch-shape variants: bit-exact.
"""
import numpy as np
import pytest

import fwi2d_amd as F
import oracle_bind as ob
from cases import layered, overthrust, rel_l2

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-5
IMG_TOL = 1e-4


def f32(x):
    return float(np.float32(x))


class Survey:
    """A small survey run through both the engine and the oracle."""

    def __init__(self, v, dh, nt, dt, fmax, xsrc, zsrc, xrec, zrec, abc_nb=0, abc_factor=0.0,
                 rbc_nb=0, varVp=0.0, varVp_rel=0.0, seed=0, max_batch=1):
        self.v = np.ascontiguousarray(v, np.float32)
        self.nx, self.nz = v.shape
        self.dh, self.dt, self.nt = f32(dh), f32(dt), nt
        self.w, self.tlag = F.ricker_wavelet(nt, dt, fmax)
        self.xsrc, self.zsrc = np.float32(xsrc), np.float32(zsrc)
        self.xrec, self.zrec = np.float32(xrec), np.float32(zrec)
        self.abc_nb, self.abc_factor, self.rbc_nb = abc_nb, abc_factor, rbc_nb
        self.varVp, self.varVp_rel, self.seed = varVp, varVp_rel, seed
        self.eng = F.Engine(self.nz, self.nx, dh, nt, dt, abc_nb, abc_factor, rbc_nb, varVp,
                            varVp_rel, seed, max_batch)
        self.eng.set_wavelet(self.w, self.tlag)
        self.eng.set_geometry(self.xsrc, self.zsrc, self.xrec, self.zrec)
        self.eng.set_model(self.v)

    # ---- oracle side
    def idx(self, nb):
        gi = lambda x: ob.grid_index(float(x), self.dh, nb)
        siz = np.array([gi(z) for z in self.zsrc], np.int32)
        six = np.array([gi(x) for x in self.xsrc], np.int32)
        riz = np.array([gi(z) for z in self.zrec], np.int32)
        rix = np.array([gi(x) for x in self.xrec], np.int32)
        return siz, six, riz, rix

    def rb_oracle(self, shot, iteration=0):
        vpad = ob.expand(self.v, self.rbc_nb)
        one = np.float32(1.0)
        vmin, vmax = self.v.min(), self.v.max()
        rel, ab = np.float32(self.varVp_rel), np.float32(self.varVp)
        vlo = vmin * (one - rel) - ab
        vhi = vmax * (one + rel) + ab
        vcfl = (np.float32(0.999) * np.float32(0.5546) * np.float32(self.dh)) / np.float32(self.dt)
        vhi = min(vhi, vcfl)
        stream = (int(iteration) << 32) | int(shot)
        return ob.rb_velocity(vpad, self.rbc_nb, ab, rel, vlo, vhi, self.seed, stream)

    def oracle_forward(self, shot, boundary, vpad=None, model=None, precision="f64"):
        nb = self.abc_nb if boundary == F.ABC else self.rbc_nb
        siz, six, riz, rix = self.idx(nb)
        if vpad is None:
            vpad = ob.expand(self.v if model is None else model, nb)
        b1d = ob.abc_b1d(nb, self.abc_factor) if boundary == F.ABC else None
        w = self.w.astype(np.float64 if precision == "f64" else np.float32)
        return ob.forward(vpad, self.dh, self.dt, self.nt, self.tlag, w,
                          int(siz[shot]), int(six[shot]), riz, rix, abc=(boundary == F.ABC),
                          b1d=b1d, nb=nb, precision=precision)


def overthrust_survey(nt=1001, nsrc=3, **kw):
    v = overthrust()
    xs = np.linspace(1000, 4000, nsrc)
    xr = np.linspace(0, 5000, 51)
    return Survey(v, 25.0, nt, 1e-3, 30.0, xs, np.full(nsrc, 100.0), xr, np.zeros(51), **kw)


def test_forward_abc_matches_oracle():
    s = overthrust_survey(abc_nb=51, abc_factor=0.004, max_batch=2)
    seis = s.eng.modeling(F.ABC)
    for k in range(3):
        ref, _, _ = s.oracle_forward(k, F.ABC)
        err = rel_l2(seis[k], ref)
        print(f"shot {k} relL2 {err:.3e}")
        assert err <= SEIS_TOL
        assert np.all(seis[k][:, 0] == 0.0)


def test_rb_velocity_bit_exact():
    s = overthrust_survey(rbc_nb=41, varVp=1000.0, seed=1234)
    for shot, it in [(0, 0), (2, 0), (1, 7)]:
        gpu = s.eng.rb_velocity(shot, it)
        ref = s.rb_oracle(shot, it)
        assert gpu.shape == ref.shape
        assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32)), (shot, it)
    # different streams differ, interior untouched
    a, b = s.eng.rb_velocity(0, 0), s.eng.rb_velocity(0, 1)
    nb = 41
    assert not np.array_equal(a, b)
    assert np.array_equal(a[nb:-nb, nb:-nb], s.v)


def test_forward_rbc_matches_oracle():
    s = overthrust_survey(rbc_nb=41, varVp=1000.0, seed=99, max_batch=3)
    seis = s.eng.modeling(F.RBC)
    for k in range(3):
        ref, _, _ = s.oracle_forward(k, F.RBC, vpad=s.rb_oracle(k))
        err = rel_l2(seis[k], ref)
        print(f"shot {k} relL2 {err:.3e}")
        assert err <= SEIS_TOL


SHAPES = [(1, 1, 8, 0), (1, 2, 17, 1), (1, 4, 64, 1), (1, 4, 1000, 0), (1, 1, 3, 1),
          (2, 1, 8, 0), (2, 2, 17, 1), (2, 4, 64, 1), (2, 4, 1000, 0), (2, 1, 3, 1), (2, 2, 5, 0)]
# three steps per launch: the window of a 144-column chunk (160 columns) is
# the largest that is staged in LDS
SHAPES3 = [(3, 1, 8, 0), (3, 2, 17, 1), (3, 4, 64, 1), (3, 4, 1000, 0), (3, 1, 3, 1),
           (3, 2, 5, 0), (3, 2, 144, 1), (3, 1, 145, 0)]
# four steps per launch: the window of a 136-column chunk (160 columns) is
# the largest that is staged in LDS
SHAPES4 = [(4, 1, 8, 0), (4, 2, 17, 1), (4, 4, 64, 1), (4, 4, 1000, 0), (4, 2, 130, 1),
           (4, 1, 3, 0), (4, 2, 136, 1), (4, 1, 137, 0)]
# the forward and reverse kernels on 64 x 64-cell LDS tiles (vz 16), one to
# four steps, xw the tile columns per wave
TILE = 16
SHAPEST = [(1, TILE, 16, 0), (2, TILE, 8, 1), (3, TILE, 4, 0), (4, TILE, 16, 1), (4, TILE, 4, 0)]


def set_shape(eng, st, vz, xw, g):
    """One launch shape for every kernel; the three- and four-step reverse
    kernels have no four-rows-per-lane form and then take two steps."""
    if st >= 3 and vz == 4:
        eng.set_tuning(vz, xw, g, kernel=1, steps=2)
        for k in (0, 2):
            eng.set_tuning(vz, xw, g, kernel=k, steps=st)
    else:
        eng.set_tuning(vz, xw, g, steps=st)


def test_launch_shapes_bit_identical():
    # every (steps per launch, rows per lane, chunk width, graph) gives the same bits
    s = overthrust_survey(abc_nb=51, abc_factor=0.004, max_batch=3)
    s.eng.set_tuning(steps=1, vz=4, xw=32, use_graph=0)
    base = s.eng.modeling(F.ABC)
    for st, vz, xw, g in SHAPES + SHAPES3 + SHAPES4 + SHAPEST:
        set_shape(s.eng, st, vz, xw, g)
        out = s.eng.modeling(F.ABC)
        assert np.array_equal(out, base), (st, vz, xw, g)


@pytest.mark.parametrize("nt", [998, 999, 1000, 1001])
def test_three_step_remainders_bit_identical(nt):
    # passes of T = nt + tlag steps with T mod 3 = 0, 1, 2 (and T mod 4 = 0-3
    # for the four-step forward) end in shorter launches; traces, image,
    # illumination, misfit and gradient equal the one-step kernels'
    s = overthrust_survey(nt=nt, nsrc=2, abc_nb=51, abc_factor=0.004, rbc_nb=41,
                          varVp=700.0, seed=4, max_batch=2)
    s.eng.set_tuning(2, 64, 0, steps=1)

    def run():
        obs = s.eng.modeling(F.ABC)
        out = [obs, s.eng.modeling(F.RBC)]
        s.eng.set_observed(obs)
        s.eng.set_model((s.v * 0.98).astype(np.float32))
        s.eng.reset()
        s.eng.migration()
        out += list(s.eng.get_image()[1:])
        s.eng.reset()
        out.append(np.float64(s.eng.gradient(iteration=1)))
        out += list(s.eng.get_gradient(100.0)[1:])
        s.eng.set_model(s.v)
        return out

    base = run()
    for st, vz, xw, g in [(3, 2, 64, 1), (3, 1, 16, 0), (3, 4, 96, 1), (4, 2, 64, 1),
                          (4, 4, 96, 0), (4, TILE, 8, 1), (3, TILE, 16, 0)]:
        set_shape(s.eng, st, vz, xw, g)
        for k, (a, b) in enumerate(zip(run(), base)):
            assert np.array_equal(a, b), (nt, vz, xw, g, k)


def test_reverse_launch_shapes_bit_identical():
    s = overthrust_survey(nsrc=2, abc_nb=51, abc_factor=0.004, rbc_nb=41, varVp=900.0,
                          seed=8, max_batch=2)
    obs = s.eng.modeling(F.ABC)
    s.eng.set_observed(obs)
    s.eng.set_model((s.v * 0.98).astype(np.float32))
    results = []
    for st, vz, xw, g in [(1, 2, 32, 0)] + SHAPES[5:] + SHAPES3 + SHAPES4 + SHAPEST:
        set_shape(s.eng, st, vz, xw, g)
        s.eng.reset()
        s.eng.migration()
        img = s.eng.get_image()
        s.eng.reset()
        f = s.eng.gradient(iteration=2)
        grd = s.eng.get_gradient(100.0)
        results.append((img[1], img[2], grd[1], grd[2], f))
    for r in results[1:]:
        for a, b in zip(r[:4], results[0][:4]):
            assert np.array_equal(a, b)
        assert r[4] == results[0][4]


def test_graph_replays_match_eager():
    # each pass is captured once into a graph and replayed afterwards; every
    # replay, with eager host work and synchronisations between, gives the
    # bits of the eager launches, for full (2-shot) and ragged (1-shot)
    # batches and for a random boundary drawn per iteration
    s = overthrust_survey(nsrc=3, abc_nb=51, abc_factor=0.004, rbc_nb=41, varVp=900.0,
                          seed=6, max_batch=2)

    def run(iteration):
        obs = s.eng.modeling(F.ABC)
        s.eng.set_observed(obs)
        s.eng.set_model((s.v * 0.98).astype(np.float32))
        s.eng.reset()
        s.eng.migration()
        img = s.eng.get_image()
        s.eng.reset()
        f = s.eng.gradient(iteration=iteration)
        grd = s.eng.get_gradient(100.0)
        s.eng.set_model(s.v)
        return [obs, img[1], img[2], grd[1], grd[2], np.float64(f)]

    s.eng.set_tuning(use_graph=0)
    base = {it: run(it) for it in (1, 2)}
    assert not np.array_equal(base[1][3], base[2][3])   # the boundary differs per iteration
    s.eng.set_tuning(use_graph=1)
    for rep, it in enumerate((1, 2, 1, 2)):
        for k, (a, b) in enumerate(zip(run(it), base[it])):
            assert np.array_equal(a, b), (rep, it, k)
        s.eng.synchronize()


def test_batching_bit_identical():
    a = overthrust_survey(nsrc=5, rbc_nb=41, varVp=800.0, seed=5, max_batch=5)
    b = overthrust_survey(nsrc=5, rbc_nb=41, varVp=800.0, seed=5, max_batch=2)
    sa = a.eng.modeling(F.RBC)
    sb = b.eng.modeling(F.RBC)
    assert np.array_equal(sa, sb)
    s3 = b.eng.modeling(F.RBC, shot0=3, nshots=1)
    assert np.array_equal(s3[0], sa[3])


def test_autotune_with_a_wider_sponge_pad():
    """The reference's own parameters.txt pads the sponge grid wider than the
    random-boundary grid (abc 1250 m -> 51 cells, rbc 1000 m -> 41 cells at
    25 m).  The autotuner runs the kernels of both grids and then zeroes what
    they touched, each buffer to its own grid's extent (every fill is checked
    against its allocation); the passes after it give the bits of the passes
    before it."""
    s = overthrust_survey(nsrc=4, abc_nb=51, abc_factor=0.004, rbc_nb=41, varVp=1000.0,
                          seed=12, max_batch=4)
    obs = s.eng.modeling(F.ABC)
    s.eng.set_observed(obs)
    s.eng.set_model((s.v * 0.98).astype(np.float32))

    def passes():
        s.eng.reset()
        s.eng.migration()
        img = s.eng.get_image()
        s.eng.reset()
        f = s.eng.gradient(iteration=1)
        grd = s.eng.get_gradient(100.0)
        return [img[1], img[2], grd[1], grd[2], np.float64(f), s.eng.modeling(F.ABC)]

    before = passes()
    fwd_us, rev_us = s.eng.autotune(4, reps=1)
    assert fwd_us > 0 and rev_us > 0
    after = passes()
    for k, (a, b) in enumerate(zip(after, before)):
        assert np.array_equal(a, b), k


def migration_oracle(s, obs):
    nb = s.rbc_nb
    siz, six, riz, rix = s.idx(nb)
    nxx, nzz = s.nx + 2 * nb, s.nz + 2 * nb
    acc = np.zeros((nxx, nzz))
    ill = np.zeros((nxx, nzz))
    for k in range(len(s.xsrc)):
        vrb = s.rb_oracle(k)
        _, P, Q = s.oracle_forward(k, F.RBC, vpad=vrb)
        ob.reverse(vrb, s.dh, s.dt, s.nt, s.tlag, P, Q, riz, rix,
                   obs[k].astype(np.float64), 0, acc=acc, illum=ill)
    crop = lambda a: np.ascontiguousarray(a[nb:nb + s.nx, nb:nb + s.nz])
    img, il = crop(acc), crop(ill)
    return ob.image_post(img, il, s.dh), img, il


def test_migration_matches_oracle():
    s = overthrust_survey(nsrc=2, abc_nb=51, abc_factor=0.004, rbc_nb=41, varVp=1000.0,
                          seed=3, max_batch=2)
    obs = np.stack([s.oracle_forward(k, F.ABC)[0] for k in range(2)]).astype(np.float32)
    s.eng.set_observed(obs)
    s.eng.reset()
    s.eng.migration()
    img, raw, ill = s.eng.get_image()
    ref, rraw, rill = migration_oracle(s, obs)
    e1, e2, e3 = rel_l2(raw, rraw), rel_l2(ill, rill), rel_l2(img, ref)
    print(f"raw image {e1:.3e} illum {e2:.3e} image {e3:.3e}")
    assert e1 <= IMG_TOL and e2 <= IMG_TOL and e3 <= IMG_TOL


def test_dense_receiver_column_matches_oracle():
    """260 receivers in one grid column (several per cell) plus a surface line:
    windows holding that column exceed the kernels' per-wave receiver table
    (256 entries) and take the global-memory path; the others are staged."""
    v = overthrust()
    xr = np.concatenate([np.linspace(0, 5000, 51), np.full(260, 2500.0)])
    zr = np.concatenate([np.zeros(51), np.linspace(0, 1900, 260)])
    s = Survey(v, 25.0, 801, 1e-3, 30.0, [1500.0, 3500.0], [100.0, 100.0], xr, zr,
               abc_nb=51, abc_factor=0.004, rbc_nb=41, varVp=1000.0, seed=3, max_batch=2)
    seis = s.eng.modeling(F.ABC)
    obs = []
    for k in range(2):
        ref, _, _ = s.oracle_forward(k, F.ABC)
        err = rel_l2(seis[k], ref)
        print(f"shot {k} seismogram relL2 {err:.3e}")
        assert err <= SEIS_TOL
        obs.append(ref)
    for st in (1, 2, 3):                    # windows wider than the table: global path
        set_shape(s.eng, st, 2, 200, 0)
        assert np.array_equal(s.eng.modeling(F.ABC), seis), st
    for st in (4, 1):                       # LDS tiles (own columns' receivers)
        set_shape(s.eng, st, TILE, 8, 0)
        assert np.array_equal(s.eng.modeling(F.ABC), seis), (st, TILE)
    obs = np.stack(obs).astype(np.float32)
    s.eng.set_observed(obs)
    rref, rraw, rill = migration_oracle(s, obs)
    first = None
    # the LDS tiles: the tiles holding the dense column stage 256 of its
    # receivers' data and read the rest from global memory
    for st, vz, xw in [(2, 2, 64), (1, 1, 8), (2, 4, 128), (1, 2, 200), (3, 2, 64), (3, 4, 200),
                       (4, TILE, 8), (1, TILE, 4), (3, TILE, 16)]:
        set_shape(s.eng, st, vz, xw, 0)
        s.eng.reset()
        s.eng.migration()
        img, raw, ill = s.eng.get_image()
        if first is None:
            e1, e2, e3 = rel_l2(raw, rraw), rel_l2(ill, rill), rel_l2(img, rref)
            print(f"raw image {e1:.3e} illum {e2:.3e} image {e3:.3e}")
            assert e1 <= IMG_TOL and e2 <= IMG_TOL and e3 <= IMG_TOL
            first = raw
        else:
            assert np.array_equal(raw, first), (st, vz, xw)


def gradient_oracle(s, obs, model, iteration, zmask, precision="f64"):
    """FP64 oracle of the FWI gradient pass and its post-processing;
    precision="f32" runs the oracle's FP32 restatement instead (the
    reference's own arithmetic order in single precision)."""
    nb = s.rbc_nb
    siz, six, riz, rix = s.idx(nb)
    nxx, nzz = s.nx + 2 * nb, s.nz + 2 * nb
    dtype = np.float64 if precision == "f64" else np.float32
    acc = np.zeros((nxx, nzz), dtype)
    ill = np.zeros((nxx, nzz), dtype)
    sumsq = 0.0
    for k in range(len(s.xsrc)):
        cal, _, _ = s.oracle_forward(k, F.ABC, model=model, precision=precision)
        res = obs[k].astype(dtype) - cal
        sumsq += float(np.sum(res.astype(np.float64) ** 2))
        vrb = s.rb_oracle(k, iteration)
        _, P, Q = s.oracle_forward(k, F.RBC, vpad=vrb, precision=precision)
        ob.reverse(vrb, s.dh, s.dt, s.nt, s.tlag, P, Q, riz, rix, res, 1, acc=acc, illum=ill,
                   precision=precision)
    crop = lambda a: np.ascontiguousarray(a[nb:nb + s.nx, nb:nb + s.nz], np.float64)
    g, il = crop(acc), crop(ill)
    return ob.gradient_post(g, il, s.dh, zmask), g, il, sumsq


def test_gradient_and_misfit_match_oracle():
    s = overthrust_survey(nsrc=2, abc_nb=51, abc_factor=0.004, rbc_nb=41, varVp=1000.0,
                          seed=11, max_batch=2)
    true_v = s.v.copy()
    obs = np.stack([s.oracle_forward(k, F.ABC)[0] for k in range(2)]).astype(np.float32)
    init = (true_v * 0.97).astype(np.float32)          # a perturbed starting model
    s.v = init
    s.eng.set_model(init)
    s.eng.set_observed(obs)
    s.eng.reset()
    sumsq = s.eng.gradient(iteration=4)
    g, raw, ill, gmax = s.eng.get_gradient(400.0)
    rg, rraw, rill, rsum = gradient_oracle(s, obs, init, 4, 400.0)
    e0 = abs(sumsq - rsum) / rsum
    e1, e2, e3 = rel_l2(raw, rraw), rel_l2(ill, rill), rel_l2(g, rg)
    print(f"misfit {e0:.3e} raw grad {e1:.3e} illum {e2:.3e} direction {e3:.3e}")
    assert e0 <= 1e-5
    assert e1 <= IMG_TOL and e2 <= IMG_TOL and e3 <= IMG_TOL
    # misfit alone agrees with the gradient pass's misfit
    assert abs(s.eng.misfit() - sumsq) <= 1e-12 * sumsq


def test_config1_forward_full_size():
    """121 x 401 at 10 m, 100-cell pad, nt=2001, dt=1 ms, 15 Hz peak (fmax 45)."""
    v = layered(121, 401, [1500, 2000, 2500, 3000], anomaly=(15, 250.0), seed=1)
    xr = np.arange(401) * 10.0
    for kw, bnd in [(dict(abc_nb=100, abc_factor=0.004), F.ABC),
                    (dict(rbc_nb=100, varVp_rel=0.4, seed=2024), F.RBC)]:
        s = Survey(v, 10.0, 2001, 1e-3, 45.0, [2000.0], [20.0], xr, np.full(401, 10.0), **kw)
        seis = s.eng.modeling(bnd)
        vpad = s.rb_oracle(0) if bnd == F.RBC else None
        ref, _, _ = s.oracle_forward(0, bnd, vpad=vpad)
        err = rel_l2(seis[0], ref)
        print(f"boundary {bnd} relL2 {err:.3e}")
        assert err <= SEIS_TOL
        # the LDS-tile forward kernel (the single-shot shape) gives the same bits
        for st, cpw in ((4, 8), (3, 16), (4, 4)):
            for k in (0, 2):
                s.eng.set_tuning(TILE, cpw, 1, kernel=k, steps=st)
            assert np.array_equal(s.eng.modeling(bnd), seis), (bnd, st, cpw)


def test_errors_are_reported():
    v = overthrust()
    eng = F.Engine(81, 201, 25.0, 101, 1e-3, abc_nb=51, abc_factor=0.004, max_batch=1)
    with pytest.raises(F.FwiError) as e:
        eng.modeling(F.ABC)
    assert e.value.code == 2          # state: nothing set yet
    w, tl = F.ricker_wavelet(101, 1e-3, 30.0)
    eng.set_wavelet(w, tl)
    with pytest.raises(F.FwiError):   # receiver outside the padded grid
        eng.set_geometry([100.0], [100.0], [1e6], [0.0])
    with pytest.raises(F.FwiError) as e:   # CFL: 20 km/s at dt/dh = 1/25 ms/m
        eng.set_model(np.full((201, 81), 20000.0, np.float32))
    assert e.value.code == 5
    with pytest.raises(F.FwiError) as e:
        eng.set_model(np.full((201, 81), -1.0, np.float32))
    assert e.value.code == 5
    eng.set_tuning(1, 64, kernel=-1, steps=3)
    for kernel in (-1, 1):            # no three-step reverse kernel with 4 rows per lane
        with pytest.raises(F.FwiError):
            eng.set_tuning(4, kernel=kernel)
    assert eng.get_tuning(1)[:2] == (3, 1)      # a rejected call changes nothing
    eng.set_tuning(4, kernel=2)
    assert eng.get_tuning(2)[:2] == (3, 4)
    for kernel in (0, 1, 2):
        with pytest.raises(F.FwiError):   # at most four steps per launch
            eng.set_tuning(kernel=kernel, steps=5)
    with pytest.raises(F.FwiError):       # no four-step reverse kernel with 4 rows per lane
        eng.set_tuning(4, kernel=1, steps=4)
    eng.set_tuning(kernel=-1, steps=4)
    assert [eng.get_tuning(k)[:2] for k in (0, 1, 2)] == [(4, 1), (4, 1), (4, 4)]
    with pytest.raises(F.FwiError):
        eng.set_tuning(8, kernel=0)
    for kernel in (-1, 0, 1):             # tile columns per wave: 4, 8 or 16
        with pytest.raises(F.FwiError):
            eng.set_tuning(TILE, 64, kernel=kernel)
    eng.set_tuning(TILE, 8, kernel=0)
    assert eng.get_tuning(0)[:3] == (4, TILE, 8) and eng.get_tuning(1)[:2] == (4, 1)
    eng.set_tuning(TILE, 4, kernel=1)
    assert eng.get_tuning(1)[:3] == (4, TILE, 4)


def test_rejected_geometry_keeps_the_previous_one():
    """A geometry that one grid accepts and the other rejects (a receiver
    700 m left of the model: inside the 60-cell sponge pad, outside the
    20-cell random pad) is refused as a whole: the previous geometry stays in
    force for both grids and gives the same seismograms and image."""
    v = overthrust()
    eng = F.Engine(81, 201, 25.0, 301, 1e-3, abc_nb=60, abc_factor=0.004, rbc_nb=20,
                   rbc_varVp_rel=0.4, rbc_seed=3, max_batch=2)
    w, tl = F.ricker_wavelet(301, 1e-3, 30.0)
    eng.set_wavelet(w, tl)
    xr = np.linspace(0.0, 5000.0, 51)
    eng.set_geometry([1500.0, 3500.0], [100.0, 100.0], xr, np.zeros(51))
    eng.set_model(v)
    seis = eng.modeling(F.ABC)
    eng.set_observed(seis)
    eng.reset()
    eng.migration()
    _, img, ill = eng.get_image()
    with pytest.raises(F.FwiError) as e:
        eng.set_geometry([1500.0], [100.0], [-700.0, 100.0], [0.0, 0.0])
    assert "receiver 0" in str(e.value)
    assert np.array_equal(eng.modeling(F.ABC), seis)
    eng.reset()
    eng.migration()
    _, img2, ill2 = eng.get_image()
    assert np.array_equal(img2, img) and np.array_equal(ill2, ill)
    eng.close()


def test_gradient_predicts_finite_difference_misfit():
    """The reference's gradient (grad += dt R v^2 lap S, inversion.cu:442) is
    the adjoint-state derivative of the misfit with respect to m = 1/v^2, up
    to a constant factor and the approximations of the reference's reverse
    pass (random-boundary source reversal, D12/D13 pairing).  Central
    differences of the misfit along localized perturbations dm below the
    sources agree with <raw gradient, dm>: the same sign and one common
    factor for every direction (within the stated spread)."""
    s = overthrust_survey(nsrc=4, abc_nb=51, abc_factor=0.004, rbc_nb=41, varVp=1000.0,
                          seed=11, max_batch=4)
    obs = s.eng.modeling(F.ABC)
    init = (s.v * 0.97).astype(np.float32)
    s.eng.set_model(init)
    s.eng.set_observed(obs)
    s.eng.reset()
    s.eng.gradient(iteration=1)
    _, raw, _, _ = s.eng.get_gradient(0.0)
    m0 = 1.0 / init.astype(np.float64) ** 2
    ix, iz = np.meshgrid(np.arange(s.nx), np.arange(s.nz), indexing="ij")

    def misfit(m):
        s.eng.set_model((1.0 / np.sqrt(m)).astype(np.float32))
        return s.eng.misfit()

    ratios = []
    for cx, cz in ((60, 50), (100, 60), (140, 45), (120, 70)):
        dm = 0.01 * m0 * np.exp(-0.5 * ((ix - cx) ** 2 + (iz - cz) ** 2) / 4.0 ** 2)
        fd = 0.5 * (misfit(m0 + dm) - misfit(m0 - dm))
        pred = float(np.sum(raw.astype(np.float64) * dm))
        ratios.append(fd / pred)
        print(f"blob at x={cx} z={cz}: finite difference {fd:.4e}, <grad, dm> {pred:.4e}, "
              f"ratio {fd / pred:.4e}")
    r = np.array(ratios)
    assert np.all(np.sign(r) == np.sign(r[0]))
    assert np.abs(r).max() / np.abs(r).min() <= 1.5


def test_full_size_invariants():
    """The 64-shot RTM survey at its real size (301 x 1001 at 10 m, 150-cell
    random pad: 64 x 601 x 1301 cells, nt 4001; 100-cell sponge grid for the
    observed data), through properties that need no oracle:
    * a shot modeled alone equals the same shot in the 64-shot batch;
    * one-, three- and four-step launch shapes (the bench's shapes among
      them) give the same raw image and illumination (both passes of the
      RTM);
    * the RTM is linear in the data: doubled data give a doubled image and
      the same illumination (which depends on the source wavefield only);
    * modeling is linear in the wavelet: a doubled wavelet gives doubled
      seismograms.
    Scaling by 2 commutes with every FP32 operation except where a value is
    subnormal: the wavefields' leading edges pass through that range, and the
    FP32 CPU port of the same algorithm shows the same departure from exact
    linearity (about 2e-6 with flush-to-zero, 5e-7 without; the FP64 oracle
    is exactly linear).  The linearity checks therefore use the seismogram
    tolerance; a source term leaking into the receiver wavefield or a stray
    scale factor would miss it by orders of magnitude."""
    from fwi2d_amd import configs as C
    s = C.CONFIG2
    true_v, mig_v = C.config2_models()
    eng = F.Engine(s.nz, s.nx, s.dh, s.nt, s.dt, abc_nb=s.abc_nb, abc_factor=s.abc_factor,
                   rbc_nb=s.nb, rbc_varVp=s.varVp, rbc_varVp_rel=s.varVp_rel, rbc_seed=s.seed,
                   max_batch=64)
    w, tlag = F.ricker_wavelet(s.nt, s.dt, s.fmax)
    eng.set_wavelet(w, tlag)
    eng.set_geometry(*s.acquisition())
    eng.set_model(true_v)
    obs = eng.modeling(F.ABC)
    assert np.isfinite(obs).all() and np.abs(obs).max() > 0
    for k in (0, 37, 63):
        alone = eng.modeling(F.ABC, shot0=k, nshots=1)
        assert np.array_equal(alone[0], obs[k]), k

    eng.set_model(mig_v)
    eng.set_observed(obs)
    eng.set_tuning(vz=4, xw=81, kernel=0, steps=3)
    eng.set_tuning(vz=2, xw=81, kernel=1, steps=3)
    eng.reset()
    eng.migration()
    _, raw3, ill3 = eng.get_image()
    eng.set_tuning(vz=2, xw=64, kernel=0, steps=1)
    eng.set_tuning(vz=1, xw=48, kernel=1, steps=1)
    eng.reset()
    eng.migration()
    _, raw1, ill1 = eng.get_image()
    assert np.abs(raw3).max() > 0
    assert np.array_equal(raw3, raw1) and np.array_equal(ill3, ill1)
    for fvz, fxw, rxw in ((2, 130, 81), (4, 130, 130)):
        eng.set_tuning(vz=fvz, xw=fxw, kernel=0, steps=4)
        eng.set_tuning(vz=2, xw=rxw, kernel=1, steps=4)
        eng.reset()
        eng.migration()
        _, raw4, ill4 = eng.get_image()
        assert np.array_equal(raw4, raw1) and np.array_equal(ill4, ill1), (fvz, fxw, rxw)

    eng.set_tuning(vz=4, xw=81, kernel=0, steps=3)
    eng.set_tuning(vz=2, xw=81, kernel=1, steps=3)
    eng.set_observed(2.0 * obs)
    eng.reset()
    eng.migration()
    _, raw2, ill2 = eng.get_image()
    e_img = rel_l2(raw2, 2.0 * raw3)
    print(f"image linearity relL2 {e_img:.3e} (equal: {np.array_equal(raw2, 2.0 * raw3)})")
    assert np.array_equal(ill2, ill3)
    assert e_img <= SEIS_TOL

    eng.set_model(true_v)
    eng.set_wavelet(2.0 * w, tlag)
    obs2 = eng.modeling(F.ABC, 0, 8)
    e_seis = rel_l2(obs2, 2.0 * obs[:8])
    print(f"seismogram linearity relL2 {e_seis:.3e} (equal: {np.array_equal(obs2, 2.0 * obs[:8])})")
    assert e_seis <= SEIS_TOL
    eng.close()


def test_one_fwi_iteration_matches_oracle():
    """One whole plain FWI iteration (inversion.cu:111-358) against the FP64
    oracle on the overthrust model, 2 shots: the gradient pass and its
    post-processing, then the C++ update (fwi_update_model: misfits of the two
    line-search models m - alpha_k g, the parabola and the step rule, and
    m - step g).  The oracle models the same two trial models from its own
    direction, and its parabola vertex (or_parabola_step) gives the step; the
    test requires the same rule, the step within 1e-3 and the model update
    m_new - m within 1e-3 (relative L2)."""
    s = overthrust_survey(nsrc=2, abc_nb=51, abc_factor=0.004, rbc_nb=41, varVp=1000.0,
                          seed=21, max_batch=2)
    from fwi2d_amd import configs as C
    true_v = s.v.copy()
    obs = np.stack([s.oracle_forward(k, F.ABC)[0] for k in range(2)]).astype(np.float32)
    init = C.smooth_slowness(true_v, 6.0)
    s.v = init
    s.eng.set_model(init)
    s.eng.set_observed(obs)
    s.eng.reset()
    zmask = 100.0
    f0 = float(np.sqrt(s.eng.gradient(iteration=1)))
    g, _, _, _ = s.eng.get_gradient(zmask)
    m = (1.0 / init.astype(np.float64) ** 2).astype(np.float32)
    dm = 1.0 / float(init.min()) ** 2 - 1.0 / float(init.max()) ** 2
    a1, a2 = 0.02 * dm, 0.05 * dm
    vstable = 0.999 * 0.5546 * s.dh / s.dt
    m_new, info = s.eng.update_model(m, g, f0, a1, a2, 1.0 / vstable ** 2)

    rg, _, _, rsum = gradient_oracle(s, obs, init, 1, zmask)
    rf = [np.sqrt(rsum)]
    m64 = 1.0 / init.astype(np.float64) ** 2
    for a in (a1, a2):
        v = 1.0 / np.sqrt(m64 - a * rg)
        rf.append(np.sqrt(sum(float(np.sum((obs[k].astype(np.float64) -
                                            s.oracle_forward(k, F.ABC, model=v)[0]) ** 2))
                              for k in range(2))))
    alpha = np.array([0.0, a1, a2])
    print(f"misfits GPU {info['f']} oracle {rf}; rule {info['rule']} step {info['step']:.6e}")
    for k in range(3):
        assert abs(info["f"][k] - rf[k]) <= 1e-5 * rf[k], k
    curv = ((rf[2] - rf[0]) / a2 - (rf[1] - rf[0]) / a1) / (a2 - a1)
    vertex = ob.parabola_step(alpha, np.array(rf))
    if curv > 0 and 0 < vertex <= 4 * a2:
        ref_rule, ref_step = 0, vertex
    else:
        k = 2 if rf[2] < rf[1] else 1
        ref_rule, ref_step = 1, alpha[k]
    print(f"oracle rule {ref_rule} step {ref_step:.6e}")
    assert info["rule"] == ref_rule
    assert abs(info["step"] - ref_step) <= 1e-3 * ref_step
    upd, rupd = m_new.astype(np.float64) - m, -ref_step * rg
    e = rel_l2(upd, rupd)
    print(f"model update relL2 {e:.3e}")
    assert e <= 1e-3
    # the engine holds the updated model on return
    held = s.eng.misfit()
    s.eng.set_model((1.0 / np.sqrt(m_new)).astype(np.float32))
    assert s.eng.misfit() == held
    s.eng.close()
