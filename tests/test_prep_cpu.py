"""CPU tests of the data-preparation helpers (fwi2d_amd/configs.py), the
restatement of the reference's Python harness: slowness smoothing
(run/generate_setup.py:43-53, scipy.ndimage.gaussian_filter), the
acquisition and parameter files (run/generate_setup.py:23-36,
run/parameters.txt) and the direct-wave mute (run/generate_input_data.py:22-46)."""
import math

import numpy as np
import pytest
import scipy.ndimage as ndi

from fwi2d_amd import configs as C


@pytest.mark.parametrize("shape,sigma", [((40, 30), 2.0), ((64, 17), 5.5), ((12, 9), 6.0)])
def test_gaussian_smooth_is_scipy_reflect(shape, sigma):
    a = np.random.default_rng(1).uniform(1500, 4500, shape)
    ref = ndi.gaussian_filter(a, sigma, mode="reflect", truncate=4.0)
    out = C.gaussian_smooth(a, sigma)
    assert np.abs(out - ref).max() <= 1e-9 * np.abs(ref).max()


def test_smooth_slowness_and_shallow_copy():
    v = C.flat_layers(40, 60, [1500.0, 2500.0, 3500.0])
    s = C.smooth_slowness(v, 4.0, keep_rows=5)
    ref = 1.0 / ndi.gaussian_filter(1.0 / v.astype(np.float64), 4.0, mode="reflect")
    assert s.dtype == np.float32
    assert np.allclose(s[:, 5:], ref[:, 5:], rtol=1e-6)
    assert np.array_equal(s[:, :5], v[:, :5])
    # smoothing slowness, not velocity: the mean slowness of a column is kept
    col = slice(None)
    assert abs(np.mean(1.0 / s[:, 20].astype(np.float64)) -
               np.mean(1.0 / ref[:, 20])) < 1e-9


def test_mute_direct_wave_closed_form():
    nrec, nt, dt = 7, 400, 1e-3
    rng = np.random.default_rng(3)
    data = rng.standard_normal((nrec, nt)).astype(np.float32)
    xr = np.linspace(0.0, 1200.0, nrec)
    zr = np.full(nrec, 10.0)
    xs, zs, vel, delay, g = 600.0, 20.0, 3600.0, 0.18, 30.0
    out = C.mute_direct_wave(data, xs, zs, xr, zr, dt, vel, delay, gstd=g)
    d = data / np.abs(data).max()
    for r in range(nrec):
        tr = int((math.hypot(xr[r] - xs, zr[r] - zs) / vel + delay) / dt)
        t = np.arange(nt)
        w = np.where(t < tr, np.exp(-0.5 * ((t - tr) / g) ** 2), 1.0)
        assert np.allclose(out[r], d[r] * w, rtol=1e-5, atol=1e-7)
    # without normalisation only the taper applies
    out2 = C.mute_direct_wave(data, xs, zs, xr, zr, dt, vel, delay, gstd=g, normalise=False)
    assert np.allclose(out2[:, -1], data[:, -1])


def test_geometry_and_parameter_files(tmp_path):
    xs = np.array([100.0, 250.5]); zs = np.array([20.0, 20.0])
    xr = np.array([0.0, 10.0, 20.25]); zr = np.array([10.0, 10.0, 10.0])
    sps, rps = tmp_path / "sps.txt", tmp_path / "rps.txt"
    C.write_geometry(sps, rps, xs, zs, xr, zr)
    assert sps.read_text().splitlines() == ["100.00,20.00", "250.50,20.00"]
    assert rps.read_text().splitlines()[2] == "20.25,10.00"
    s = C.CONFIG2
    folders = C.ensure_dirs(str(tmp_path))
    p = tmp_path / "parameters.txt"
    C.write_parameters(p, s, "model.bin", sps, rps, folders)
    kv = dict(line.split(" = ", 1) for line in p.read_text().splitlines())
    # the reference derives pad widths as (int)(length/dh) + 1
    for key, nb in (("abc_length", s.abc_nb), ("inv_rbc_length", s.nb), ("mig_rbc_length", s.nb)):
        assert int(float(kv[key]) / s.dh) + 1 == nb
    assert int(kv["x_samples"]) == s.nx and int(kv["z_samples"]) == s.nz
    assert float(kv["max_frequency"]) == s.fmax and int(kv["time_samples"]) == s.nt
