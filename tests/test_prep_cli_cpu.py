"""CPU tests of the data-preparation command line (`python -m fwi2d_amd.configs
setup | init | mute`), the restatement of the reference's run/generate_setup.py
and run/generate_input_data.py on the files the executables read and write.

The checks tie the three commands to the executables' file contract: the
parameter files name existing models and geometry; the muted files land
exactly where the migration's mig_input_folder + mig_input_prefix + k + ".bin"
looks; the starting model is the smoothed slowness (checked against
scipy.ndimage) with the rows above depth_mask kept."""
import os
import subprocess
import sys

import numpy as np
import pytest
import scipy.ndimage as ndi

from fwi2d_amd import configs as C
from fwi2d_amd.viewer import read_parameter, stoi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_setup_writes_every_input_of_the_three_executables(tmp_path):
    root = str(tmp_path)
    files = C.main(["setup", "forward-1shot", root, "--depth-mask", "50",
                    "--max-iteration", "7"])
    s = C.CONFIG1
    assert [os.path.basename(f) for f in files] == [
        "parameters_modeling.txt", "parameters_migration.txt", "parameters_inversion.txt"]
    mod, mig, inv = files
    true_file = read_parameter(mod, "model_file")
    init_file = read_parameter(mig, "model_file")
    assert read_parameter(inv, "model_file") == init_file
    assert init_file == true_file.replace("_true_", "_init_") and init_file != true_file
    true = np.fromfile(true_file, np.float32)
    init = np.fromfile(init_file, np.float32)
    assert true.size == init.size == s.nx * s.nz
    true, init = true.reshape(s.nx, s.nz), init.reshape(s.nx, s.nz)
    assert np.array_equal(true, C.config1_model(1))
    keep = int(50.0 / s.dh)
    assert np.array_equal(init[:, :keep], true[:, :keep])
    assert not np.array_equal(init[:, keep:], true[:, keep:])
    sps = np.loadtxt(read_parameter(mod, "SPS"), delimiter=",", ndmin=2)
    rps = np.loadtxt(read_parameter(mod, "RPS"), delimiter=",", ndmin=2)
    assert sps.shape == (s.nsrc, 2) and rps.shape == (s.nrec, 2)
    assert stoi(read_parameter(inv, "max_iteration")) == 7
    assert float(read_parameter(inv, "depth_mask")) == 50.0
    # inversion reads the modeling output; migration reads the muted copies
    prefix = C.seismogram_prefix(s.nt, s.nrec, s.dt)
    assert read_parameter(inv, "inv_input_prefix") == prefix
    assert read_parameter(mig, "mig_input_prefix") == "input_RTM_" + prefix
    assert (read_parameter(mod, "mod_output_folder") == read_parameter(inv, "inv_input_folder")
            == read_parameter(mig, "mig_input_folder"))
    for key in ("mod_output_folder", "mig_output_folder", "inv_output_folder",
                "inv_residuo_folder"):
        assert os.path.isdir(read_parameter(mod, key)), key


def test_setup_rejects_an_unknown_configuration(tmp_path):
    with pytest.raises(SystemExit):
        C.main(["setup", "no-such-config", str(tmp_path)])


def fake_modeling(params, seed=0):
    """Seismogram files under the modeling executable's names, with random
    traces; returns their arrays [shot][nrec][nt]."""
    nt = stoi(read_parameter(params, "time_samples"))
    dt = float(read_parameter(params, "time_spacing"))
    folder = read_parameter(params, "mod_output_folder")
    ns = len(np.loadtxt(read_parameter(params, "SPS"), delimiter=",", ndmin=2))
    nr = len(np.loadtxt(read_parameter(params, "RPS"), delimiter=",", ndmin=2))
    rng = np.random.default_rng(seed)
    data = rng.standard_normal((ns, nr, nt)).astype(np.float32) * rng.uniform(0.1, 9.0, (ns, 1, 1))
    prefix = C.seismogram_prefix(nt, nr, dt)
    for k in range(ns):
        data[k].astype(np.float32).tofile(f"{folder}{prefix}{k + 1}.bin")
    return data


def small_run(tmp_path, nsrc=3, nrec=40, nt=700, dt=1e-3):
    s = C.Survey("small", 40, 120, 10.0, nt, dt, 30.0, 20, nsrc, nrec, 20.0, 10.0)
    folders = C.ensure_dirs(str(tmp_path))
    true_file = os.path.join(folders["models"], "small_true_40x120_10m.bin")
    v = C.flat_layers(40, 120, [1500.0, 2200.0, 3000.0])
    C.gaussian_lens(v, 25.0, 60.0, 5.0, 400.0)
    v.tofile(true_file)
    sps, rps = str(tmp_path / "SPS.txt"), str(tmp_path / "RPS.txt")
    C.write_geometry(sps, rps, *s.acquisition())
    prefix = C.seismogram_prefix(nt, nrec, dt)
    params = str(tmp_path / "parameters.txt")
    C.write_parameters(params, s, true_file, sps, rps, folders, depth_mask=60.0,
                       mig_prefix="input_RTM_" + prefix, inv_prefix=prefix)
    return s, v, params


def test_mute_writes_the_migration_inputs(tmp_path):
    s, _, params = small_run(tmp_path)
    data = fake_modeling(params)
    written = C.main(["mute", params])
    folder = read_parameter(params, "mig_input_folder")
    prefix = read_parameter(params, "mig_input_prefix")
    assert written == [f"{folder}{prefix}{k + 1}.bin" for k in range(s.nsrc)]
    sps = np.loadtxt(read_parameter(params, "SPS"), np.float32, delimiter=",", ndmin=2)
    rps = np.loadtxt(read_parameter(params, "RPS"), np.float32, delimiter=",", ndmin=2)
    t = np.arange(s.nt)
    for k, path in enumerate(written):
        out = np.fromfile(path, np.float32).reshape(s.nrec, s.nt)
        d = data[k] / np.abs(data[k]).max()
        dist = np.hypot(rps[:, 0] - sps[k, 0], rps[:, 1] - sps[k, 1]).astype(np.float64)
        onset = (dist / 3600.0 + 0.18) / s.dt
        for r in range(s.nrec):
            # after the direct-wave onset the normalised trace is untouched
            late = t >= np.ceil(onset[r]) + 1
            assert np.allclose(out[r, late], d[r, late], rtol=1e-6, atol=0)
            # one gstd (30 samples) before the onset the taper is exp(-1/2)
            tr = int(np.floor(onset[r]))
            if 0 <= tr - 30 < s.nt and abs(onset[r] - round(onset[r])) > 1e-3:
                assert out[r, tr - 30] == pytest.approx(d[r, tr - 30] * np.exp(-0.5), rel=1e-5)
        assert np.abs(out).max() <= 1.0 + 1e-7


def test_mute_options_and_missing_input(tmp_path):
    s, _, params = small_run(tmp_path, nsrc=2)
    data = fake_modeling(params, seed=4)
    a = C.mute(params, velocity=1e9, delay=0.0, gstd=30.0)
    # infinite speed and no delay: the onset is sample 0, nothing is tapered
    for k, path in enumerate(a):
        out = np.fromfile(path, np.float32).reshape(s.nrec, s.nt)
        assert np.allclose(out, data[k] / np.abs(data[k]).max(), rtol=1e-6, atol=0)
    folder = read_parameter(params, "mod_output_folder")
    prefix = C.seismogram_prefix(s.nt, s.nrec, s.dt)
    for f in a:
        os.remove(f)
    os.remove(f"{folder}{prefix}2.bin")
    with pytest.raises(FileNotFoundError):
        C.mute(params)
    assert not any(os.path.exists(f) for f in a)       # nothing written for shot 1 either
    data[1, :, :10].tofile(f"{folder}{prefix}2.bin")    # short file
    with pytest.raises(SystemExit, match="floats"):
        C.mute(params)
    assert not any(os.path.exists(f) for f in a)


def test_initial_model_from_parameters(tmp_path):
    s, v, params = small_run(tmp_path)
    out = C.main(["init", params, "--sigma", "3", "--sources", "100,1000,4,20",
                  "--receivers", "0,1190,120,10"])[0]
    assert os.path.basename(out) == "small_init_40x120_10m.bin"
    init = np.fromfile(out, np.float32).reshape(s.nx, s.nz)
    ref = 1.0 / ndi.gaussian_filter(1.0 / v.astype(np.float64), 3.0, mode="reflect",
                                    truncate=4.0)
    keep = int(60.0 / s.dh)
    assert np.array_equal(init[:, :keep], v[:, :keep])
    assert np.allclose(init[:, keep:], ref[:, keep:], rtol=1e-6, atol=0)
    sps = np.loadtxt(read_parameter(params, "SPS"), delimiter=",", ndmin=2)
    rps = np.loadtxt(read_parameter(params, "RPS"), delimiter=",", ndmin=2)
    assert sps.tolist() == [[100.0, 20.0], [400.0, 20.0], [700.0, 20.0], [1000.0, 20.0]]
    assert rps.shape == (120, 2) and rps[-1].tolist() == [1190.0, 10.0]


def test_initial_model_refuses_to_overwrite_the_model(tmp_path):
    s, v, params = small_run(tmp_path)
    text = open(params).read().replace("small_true_", "small_")
    os.rename(read_parameter(params, "model_file"),
              read_parameter(params, "model_file").replace("small_true_", "small_"))
    open(params, "w").write(text)
    with pytest.raises(SystemExit, match="overwrite"):
        C.initial_model(params)


def test_command_line_entry_point(tmp_path):
    r = subprocess.run([sys.executable, "-m", "fwi2d_amd.configs", "setup", "forward-1shot",
                        str(tmp_path)], capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split()[0].endswith("parameters_modeling.txt")
    r = subprocess.run([sys.executable, "-m", "fwi2d_amd.configs"], capture_output=True,
                       text=True, cwd=ROOT, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_initial_model_checks_its_inputs_before_writing(tmp_path):
    """A short model file is refused before SPS/RPS are rewritten."""
    s, v, params = small_run(tmp_path)
    sps = read_parameter(params, "SPS")
    before = open(sps).read()
    v.ravel()[: v.size // 2].tofile(read_parameter(params, "model_file"))
    with pytest.raises(SystemExit, match="floats"):
        C.initial_model(params, sources="0,100,3,20")
    assert open(sps).read() == before
    v.tofile(read_parameter(params, "model_file"))
    with pytest.raises(ValueError):
        C.initial_model(params, sources="0,100,20")       # malformed line spec
    assert open(sps).read() == before
