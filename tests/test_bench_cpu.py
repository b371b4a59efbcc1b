"""CPU tests of bench.py: it measures one GPU only and says so when asked for
more, and it refuses unknown workloads, both before anything touches a GPU or
loads the library; its CPU-baseline leg runs and describes its sample."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, cwd=ROOT, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    for n in ("2", "8"):
        r = run("--gpus", n, "--steps", "1", "--warmup", "0")
        assert r.returncode == 2
        assert r.stdout == ""                       # no JSON line is printed
        assert f"--gpus {n}" in r.stderr and "one GPU" in r.stderr


def test_unknown_workload_is_refused():
    r = run("--workload", "no-such-workload")
    assert r.returncode == 2 and "invalid choice" in r.stderr and r.stdout == ""


def test_cpu_baseline_leg_on_a_tiny_workload():
    # the reported CPU baseline (FP32 OpenMP statement of the reference
    # stencil) runs the same RTM per shot and states its sample and threads
    sys.path.insert(0, ROOT)
    import numpy as np
    import bench
    import fwi2d_amd as F
    from fwi2d_amd import configs as C
    s = C.Survey("tiny", 24, 40, 10.0, 80, 1e-3, 30.0, 12, 2, 8, 20.0, 10.0)
    v = C.flat_layers(s.nz, s.nx, [1500.0, 2000.0])
    assert v.shape == (s.nx, s.nz)
    xs, zs, xr, zr = s.acquisition()
    w, tlag = F.ricker_wavelet(s.nt, s.dt, s.fmax)
    obs = np.random.default_rng(0).standard_normal((s.nsrc, s.nrec, s.nt)).astype(np.float32)
    out = bench.cpu_baseline(s, v, xs, zs, xr, zr, w, tlag, obs, 2, 2)
    assert out["kind"] == "port" and out["cores"] == 2 and out["unit"] == "Gcell-updates/s"
    assert out["value"] > 0
    assert out["sample"].startswith("2 of 2 shots")
    assert f"{s.nz + 2 * s.nb}x{s.nx + 2 * s.nb} cells x {s.nt + tlag} steps" in out["sample"]


def test_roofline_record_is_a_fraction_of_the_hbm_peak():
    # the reverse kernel's record from its event profile (3-step launches,
    # 64 shots of 601 x 1301 cells): before the counter passes its bytes are
    # the compulsory 52 B per cell and launch; after them, the measured
    # traffic.  The per-step-equivalent figure (22 B per cell-update of the
    # three steps) stays under its own key and may exceed 1.
    sys.path.insert(0, ROOT)
    import bench
    cells = 64 * 601 * 1301
    n = 100
    ms = n * 0.7095                              # 709.5 us per launch
    r = bench.roof(ms, n, n * cells * 3 * 2.0, bench.BYTES_REV, 5400.0, kind=1, steps=3)
    assert r["compulsory_bytes_per_launch"] == cells * 52.0
    assert abs(r["frac"] - cells * 52.0 / 709.5e-6 / 8e12) < 1e-4
    pse = r["per_step_equivalent"]
    assert pse["algorithmic_bytes_per_launch"] == cells * 3 * 2 * 22.0
    assert pse["frac"] > 1.0 > r["frac"]
    bench.set_traffic(r, 1.894e9, 1.246e9)
    assert r["traffic"] == 3.140e9
    assert abs(r["frac"] - 3.140e9 / 709.5e-6 / 8e12) < 1e-4 and r["frac"] <= 1.0
    assert abs(r["achieved"] - r["frac"] * 8000.0) < 0.1
    assert abs(r["traffic_over_compulsory"] - 3.140e9 / (cells * 52.0)) < 1e-4


def test_counter_kernel_names_match_the_built_kernels():
    """The counter passes find their kernels by name fragment (kernel_key):
    every launch shape the autotuner can choose for the random-boundary
    forward and the imaging reverse names a kernel the library holds."""
    import shutil
    import pytest
    sys.path.insert(0, ROOT)
    import bench
    lib = os.path.join(ROOT, "fwi2d_amd", "lib", "libfwi2d.so")
    nm = shutil.which("nm")
    if not os.path.exists(lib) or nm is None:
        pytest.skip("library not built or nm missing")
    names = subprocess.run([nm, "-C", lib], capture_output=True, text=True).stdout
    for steps in (1, 2, 3, 4):
        assert bench.kernel_key(0, steps, 16) in names, (0, steps, 16)
        for vz in (1, 2, 4):
            assert bench.kernel_key(0, steps, vz) in names, (0, steps, vz)
            if not (steps >= 3 and vz == 4):
                assert bench.kernel_key(1, steps, vz) in names, (1, steps, vz)
