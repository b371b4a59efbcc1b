"""GPU tests of the learned-preconditioner FWI (fwi2d_amd.learned):

* a network with zero weights gives the control direction -g, so the
  learned run equals the control run step for step;
* an ascent direction (+g) never passes the acceptance rule, so the run
  equals plain FWI step for step ("never worse than the plain step");
* the trained networks (tests/golden/preconditioner_48models.bin,
  preconditioner_mid48.bin and preconditioner_mid96.bin, made by
  tests/golden/make_preconditioner.py) gain over plain FWI on a held-out
  model of the comparison family at their training size (misfit and model
  error; misfit for the 96-model network);
* sample generation and training resume from their checkpoints.
"""
import dataclasses
import os

import numpy as np
import pytest

from fwi2d_amd import abi as A
from fwi2d_amd import configs as C
from fwi2d_amd import learned as L

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SMALL = dataclasses.replace(L.TRAIN_SMALL, nx=256, nsrc=12, nrec=128, seed=9,
                            band_iters=(2, 2, 2))


def survey(seed):
    v, v0 = C.config5_models(seed, SMALL.nz, SMALL.nx)
    return L.Fwi(SMALL, v, v0)


def trace(hist):
    return [(h["misfit"], h["step"], h["taken"]) for h in hist]


def test_zero_network_equals_control():
    fwi = survey(20001)
    n = A.Net(1, SMALL.nx, SMALL.nz, 0).nparams
    d = L.Direction(SMALL.nx, SMALL.nz, np.zeros(n, np.float32))
    a, ma = L.run_fwi(fwi, d, band_iters=(1, 1, 1))
    b, mb = L.run_fwi(fwi, L.plain_direction, band_iters=(1, 1, 1))
    d.close()
    fwi.close()
    assert trace(a) == trace(b)
    assert np.array_equal(ma, mb)


def test_ascent_direction_never_beats_the_plain_step():
    fwi = survey(20002)
    up, _ = L.run_fwi(fwi, lambda raw, ill, m, g: g, band_iters=(1, 1, 1))
    plain, _ = L.run_fwi(fwi, None, band_iters=(1, 1, 1))
    fwi.close()
    assert all(h["taken"] in ("plain", "end") for h in up)
    assert [h["misfit"] for h in up] == [h["misfit"] for h in plain]


def test_trained_network_gains_over_plain_fwi():
    """Held-out model of the comparison family at the training grid size
    (128 x 384, 32 shots), bands 4/8/12 Hz, 3 iterations each: the learned
    run (trained network + acceptance rule) ends with a lower misfit and a
    lower model error than plain FWI.  The control (same rule along -g) is
    printed beside it: it separates the gain of the longer accepted steps
    from that of the network's direction."""
    params = np.fromfile(os.path.join(GOLDEN, "preconditioner_48models.bin"), np.float32)
    s = dataclasses.replace(L.COMPARE_SMALL, band_iters=(3, 3, 3))
    v, v0 = C.config5_models(30003, s.nz, s.nx)
    fwi = L.Fwi(s, v, v0)
    d = L.Direction(s.nx, s.nz, params)
    runs = {"plain": L.run_fwi(fwi, None)[0], "control": L.run_fwi(fwi, L.plain_direction)[0],
            "learned": L.run_fwi(fwi, d)[0]}
    d.close()
    fwi.close()
    for k, h in runs.items():
        print(f"{k:8s} final misfit {h[-1]['misfit']:.4e}  model rms error "
              f"{h[-1]['model_rms_error']:.1f} m/s  time {h[-1]['time']:.2f} s  "
              f"trial steps {sum(e['taken'] == 'learned' for e in h)}")
    f = {k: h[-1]["misfit"] for k, h in runs.items()}
    e = {k: h[-1]["model_rms_error"] for k, h in runs.items()}
    assert f["learned"] < 0.8 * f["plain"]
    assert e["learned"] < e["plain"]


def mid_size_runs(fixture):
    """Plain and learned FWI at 200 x 800 (64 shots, bands 4/8/12 Hz, 2
    iterations each) on held-out model 30004 of the comparison family."""
    params = np.fromfile(os.path.join(GOLDEN, fixture), np.float32)
    s = dataclasses.replace(L.TRAIN_MID, seed=5, band_iters=(2, 2, 2))
    v, v0 = C.config5_models(30004, s.nz, s.nx)
    fwi = L.Fwi(s, v, v0)
    d = L.Direction(s.nx, s.nz, params)
    runs = {"plain": L.run_fwi(fwi, None)[0], "learned": L.run_fwi(fwi, d)[0]}
    d.close()
    fwi.close()
    for k, h in runs.items():
        print(f"{k:8s} final misfit {h[-1]['misfit']:.4e}  model rms error "
              f"{h[-1]['model_rms_error']:.1f} m/s  time {h[-1]['time']:.2f} s  "
              f"trial steps {sum(e['taken'] == 'learned' for e in h)}")
    f = {k: h[-1]["misfit"] for k, h in runs.items()}
    e = {k: h[-1]["model_rms_error"] for k, h in runs.items()}
    return f, e


def test_mid_trained_network_gains_over_plain_fwi():
    """The network trained at 200 x 800 on 48 models
    (tests/golden/preconditioner_mid48.bin) ends with a lower misfit and a
    lower model error than plain FWI at that size."""
    f, e = mid_size_runs("preconditioner_mid48.bin")
    assert f["learned"] < f["plain"]
    assert e["learned"] < e["plain"]


def test_mid96_network_lowers_the_misfit():
    """The network trained at 200 x 800 on 96 models
    (tests/golden/preconditioner_mid96.bin, the best at full size) ends with
    a misfit about 9 % below plain FWI's at that size; its model error there
    is 2.6 m/s above plain's (211.5 against 208.9 m/s), while at full size it
    has the lowest model error of all runs."""
    f, e = mid_size_runs("preconditioner_mid96.bin")
    assert f["learned"] < 0.95 * f["plain"]


def test_training_resumes_bit_for_bit(tmp_path):
    """Training stopped after one epoch and resumed from its checkpoint ends
    with the same parameters and loss history as a run straight through."""
    rng = np.random.default_rng(12)
    x = rng.standard_normal((20, 3, 16, 24)).astype(np.float32)
    y = rng.standard_normal((20, 16, 24)).astype(np.float32)
    kw = dict(batch=4, lr=2e-3, seed=3, flip=True)
    pa, ha, va = L.train(x, y, epochs=3, **kw)
    ck = str(tmp_path / "train.npz")
    L.train(x, y, epochs=1, checkpoint=ck, **kw)
    logs = []
    pb, hb, vb = L.train(x, y, epochs=3, checkpoint=ck, log=logs.append, **kw)
    assert logs[0] == {"resumed_from": ck, "epoch": 0}
    assert [e["epoch"] for e in logs[1:]] == [1, 2]
    assert np.array_equal(pa, pb) and ha == hb and np.array_equal(va, vb)
    # a finished checkpoint resumes to nothing left to do
    pc, hc, _ = L.train(x, y, epochs=3, checkpoint=ck, **kw)
    assert np.array_equal(pc, pa) and hc == ha
    with pytest.raises(ValueError, match="another training problem"):
        L.train(x, y, epochs=3, checkpoint=ck, **dict(kw, lr=1e-3))


TINY = dataclasses.replace(L.TRAIN_SMALL, nz=48, nx=96, nt=900, nsrc=3, nrec=48, rbc_nb=24,
                           abc_nb=24, band_iters=(1, 1, 1))


def test_sample_generation_resumes_from_its_cache(tmp_path):
    """Samples of seeds already in the cache folder are read, not recomputed,
    and equal what a fresh run gives; a cache of another setup is refused."""
    cache = str(tmp_path / "samples")
    xa, ya, ga, sa = L.make_samples(TINY, [1, 2])
    L.make_samples(TINY, [1], cache=cache)
    logs = []
    xb, yb, gb, sb = L.make_samples(TINY, [1, 2], cache=cache, log=logs.append)
    assert [e["cached"] for e in logs] == [True, False]
    assert sorted(os.listdir(cache)) == ["samples_1.npz", "samples_2.npz"]
    for a, b in ((xa, xb), (ya, yb), (ga, gb), (sa, sb)):
        assert a.shape == b.shape and np.array_equal(a, b)
    with pytest.raises(ValueError, match="another setup"):
        L.make_samples(dataclasses.replace(TINY, seed=5), [1], cache=cache)
