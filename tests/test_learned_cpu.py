"""CPU checks of the learned-FWI tools' host logic and of the seeded model
families they train and compare on."""
import numpy as np

from fwi2d_amd import configs as C
from fwi2d_amd import learned as L


def test_band_schedule_and_acquisition():
    s = L.COMPARE_SMALL
    assert s.iterations == 20
    bands = [s.band(i) for i in range(s.iterations)]
    assert bands == [0] * 7 + [1] * 7 + [2] * 6
    assert s.band(100) == 2
    xs, zs, xr, zr = s.acquisition()
    assert xs.size == s.nsrc and xr.size == s.nrec
    assert xs.min() > 0 and xs.max() < (s.nx - 1) * s.dh
    assert np.all(zs == s.zsrc) and np.all(zr == s.zrec)
    # the pad keeps the edge factor exp(-0.35^2) for any width
    assert abs(s.abc_factor * s.abc_nb - 0.35) < 1e-12


def test_normalise_inputs_per_channel():
    rng = np.random.default_rng(0)
    raw, ill, m = (rng.standard_normal((8, 12)) * k for k in (1e-7, 3e4, 2e-7))
    x = L.normalise_inputs(raw, ill, m)
    assert x.shape == (3, 8, 12) and x.dtype == np.float32
    for c, a in enumerate((raw, ill, m)):
        assert abs(np.max(np.abs(x[c])) - 1.0) < 1e-6
        assert np.allclose(x[c], a / np.max(np.abs(a)), rtol=1e-6)
    z = L.normalise_inputs(np.zeros((4, 4)), np.ones((4, 4)), np.ones((4, 4)))
    assert np.all(z[0] == 0)


def test_time_to_misfit_and_cosine():
    hist = [{"band": 0, "misfit": 1.0, "time": 1.0, "taken": "plain"},
            {"band": 1, "misfit": 5.0, "time": 2.0, "taken": "plain"},
            {"band": 1, "misfit": 3.0, "time": 3.0, "taken": "plain"},
            {"band": 1, "misfit": 2.5, "time": 3.5, "taken": "end"}]
    assert L.time_to_misfit(hist, 3.0, 1) == 3.0
    assert L.time_to_misfit(hist, 2.5, 1) == 3.5
    assert L.time_to_misfit(hist, 1.0, 1) is None      # band 0's misfit is not compared
    assert L.time_to_misfit(hist, 1.0, 0) == 1.0
    a = np.array([1.0, 2.0, 3.0])
    assert abs(L.cosine(a, 2 * a) - 1.0) < 1e-12
    assert abs(L.cosine(a, -a) + 1.0) < 1e-12


def test_model_families_are_seeded():
    v1, s1 = C.config4_models(3, 64, 192)
    v2, s2 = C.config4_models(3, 64, 192)
    v3, _ = C.config4_models(4, 64, 192)
    assert v1.shape == (192, 64) and v1.dtype == np.float32
    assert np.array_equal(v1, v2) and np.array_equal(s1, s2)
    assert not np.array_equal(v1, v3)
    assert v1.min() >= 1500.0 and v1.max() <= 4500.0
    v5, s5 = C.config5_models(5, 64, 256)
    assert v5.min() >= 1500.0 and v5.max() <= 4500.0
    assert (v5 == 4500.0).any()                      # the salt bodies
    # the starting model is smoother than the true model
    tv = lambda a: np.abs(np.diff(a, axis=1)).sum()
    assert tv(s5) < tv(v5)


def test_reports_of_separate_sessions_merge(tmp_path):
    """Runs of one comparison made in separate sessions summarise together:
    the target is plain FWI's final misfit in its last band, and each run's
    time is the first point of its trajectory at or below it."""
    import json
    a = {"runs": {"plain": [{"misfit": 9.0, "band": 1, "time": 10.0, "taken": "plain",
                             "model_rms_error": 5.0},
                            {"misfit": 4.0, "band": 1, "time": 20.0, "taken": "end",
                             "model_rms_error": 4.0}]},
         "trajectories": {"plain": [{"time": 10.0, "misfit": 9.0, "band": 1},
                                    {"time": 20.0, "misfit": 4.0, "band": 1}]}}
    b = {"runs": {"learned": [{"misfit": 9.0, "band": 1, "time": 4.0, "taken": "trial",
                               "model_rms_error": 4.5},
                              {"misfit": 3.5, "band": 1, "time": 9.0, "taken": "end",
                               "model_rms_error": 3.0}]},
         "trajectories": {"learned": [{"time": 3.0, "misfit": 9.0, "band": 1},
                                      {"time": 4.0, "misfit": 3.9, "band": 1},
                                      {"time": 9.0, "misfit": 3.5, "band": 1}]}}
    paths = []
    for name, r in (("a", a), ("b", b)):
        p = tmp_path / f"{name}.json"
        p.write_text(json.dumps(r))
        paths.append(str(p))
    s = L.main(["--merge"] + paths)
    assert s["plain_final_misfit"] == 4.0 and s["plain_time_s"] == 20.0
    assert s["runs"]["learned"]["time_to_plain_final_misfit_s"] == 4.0
    assert s["runs"]["learned"]["time_ratio"] == 0.2
    assert s["runs"]["learned"]["trial_steps_taken"] == 1
    assert s["runs"]["plain"]["time_ratio"] == 1.0
