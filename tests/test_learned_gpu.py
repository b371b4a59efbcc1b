"""GPU tests of the learned-preconditioner FWI (fwi2d_amd.learned over the
C++ model update of include/fwi2d_update.h):

* a network with zero weights gives the control direction -g, so the
  learned run equals the control run step for step;
* an ascent direction (+g) never passes the acceptance rule, so the run
  equals plain FWI step for step ("never worse than the plain step");
* the acceptance rule compares a trial with the forward-modeled misfit of
  the plain update, which the update reports and which equals the misfit
  the next gradient pass measures for that model;
* the trained networks (tests/golden/preconditioner_*.bin, made by
  tests/golden/make_preconditioner.py) against plain FWI and against the
  zero-network control under the same rule and schedule, on held-out
  models of the comparison family;
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
    return [(h["misfit"], h.get("step"), h["taken"]) for h in hist]


def test_zero_network_equals_control():
    fwi = survey(20001)
    net = L.network(SMALL.nx, SMALL.nz)
    net.set_params(np.zeros(net.nparams, np.float32))
    a, _, ma = L.run_fwi(fwi, A.TRIAL_NETWORK, net=net, band_iters=(1, 1, 1))
    b, _, mb = L.run_fwi(fwi, A.TRIAL_PLAIN, band_iters=(1, 1, 1))
    net.close()
    fwi.close()
    assert trace(a) == trace(b)
    assert np.array_equal(ma, mb)


def test_ascent_direction_never_beats_the_plain_step():
    fwi = survey(20002)
    up, _, mu = L.run_fwi(fwi, A.TRIAL_GIVEN, direction=lambda raw, ill, m, g: g,
                          band_iters=(1, 1, 1))
    plain, _, mp = L.run_fwi(fwi, A.TRIAL_NONE, band_iters=(1, 1, 1))
    fwi.close()
    assert all(h["taken"] in ("plain", "end") for h in up)
    assert [h["misfit"] for h in up] == [h["misfit"] for h in plain]
    assert np.array_equal(mu, mp)


def test_plain_misfit_is_the_next_passes_misfit():
    """The control run's update models the plain update's misfit (or takes
    it from the line search); when the plain update is kept, the next
    gradient pass measures the same model in the same band and must report
    the same misfit, bit for bit."""
    fwi = survey(20003)
    h, _, _ = L.run_fwi(fwi, A.TRIAL_PLAIN, band_iters=(2, 1, 1))
    fwi.close()
    checked = 0
    for a, b in zip(h[:-1], h[1:]):
        if a["taken"] == "plain" and a["band"] == b["band"] and a["step"] > 0:
            assert a["plain_misfit"] == b["misfit"], (a, b)
            checked += 1
        if a["taken"] == "trial" and a["band"] == b["band"]:
            assert a["trial_misfit"] == b["misfit"], (a, b)
            assert a["trial_misfit"] < a["plain_misfit"]
            checked += 1
    assert checked >= 2


def report(runs):
    for k, h in runs.items():
        print(f"{k:8s} final misfit {h[-1]['misfit']:.5e}  model rms error "
              f"{h[-1]['model_rms_error']:.2f} m/s  time {h[-1]['time']:.2f} s  "
              f"trial steps {sum(e['taken'] == 'trial' for e in h)}", flush=True)
    return ({k: h[-1]["misfit"] for k, h in runs.items()},
            {k: h[-1]["model_rms_error"] for k, h in runs.items()})


def compare(setup, fixture, model_seed, arms=("plain", "control", "learned")):
    params = np.fromfile(os.path.join(GOLDEN, fixture), np.float32)
    v, v0 = C.config5_models(model_seed, setup.nz, setup.nx)
    fwi = L.Fwi(setup, v, v0)
    net = L.network(setup.nx, setup.nz, params)
    runs = {k: L.run_fwi(fwi, L.RUNS[k], net=net)[0] for k in arms}
    net.close()
    fwi.close()
    return report(runs)


def test_trained_network_gains_over_plain_fwi():
    """Held-out model of the comparison family at the training grid size
    (128 x 384, 32 shots), bands 4/8/12 Hz, 3 iterations each: the learned
    run (trained network + acceptance rule) ends with a lower misfit and a
    lower model error than plain FWI.  The control (same rule along -g) is
    printed beside it: it separates the gain of the longer accepted steps
    from that of the network's direction."""
    s = dataclasses.replace(L.COMPARE_SMALL, band_iters=(3, 3, 3))
    f, e = compare(s, "preconditioner_48models.bin", 30003)
    assert f["learned"] < 0.8 * f["plain"]
    assert e["learned"] < e["plain"]


MID = dataclasses.replace(L.TRAIN_MID, seed=5, band_iters=(2, 2, 2))


def test_mid_trained_networks_against_plain_and_control():
    """The networks trained at 200 x 800 (48 and 96 models) on held-out model
    30004 at that size (64 shots, bands 4/8/12 Hz, 2 iterations each), against
    plain FWI and against the zero-network control under the same acceptance
    rule and schedule."""
    f48, e48 = compare(MID, "preconditioner_mid48.bin", 30004)
    f96, e96 = compare(MID, "preconditioner_mid96.bin", 30004, arms=("learned",))
    print(f"96-model network: final misfit {f96['learned']:.5e}, model error "
          f"{e96['learned']:.2f} m/s", flush=True)
    assert f48["learned"] < f48["plain"] and f96["learned"] < f48["plain"]
    assert f48["control"] < f48["plain"]


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
