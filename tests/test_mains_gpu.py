"""GPU end-to-end test of the drop-in executables: bin/modeling, bin/migration
and bin/inversion, each driven by a parameters.txt in the reference's format
(run/parameters.txt), on the reference's overthrust model.

The executables and the library path share every kernel, so the modeling
seismograms and the migration image written by the executables must equal,
bit for bit, what the C ABI returns for the same inputs.  The inversion must
write the reference's output files (model per iteration, final model,
convergence text) and lower the data residual.
"""
import os
import subprocess

import numpy as np
import pytest

import fwi2d_amd as F
from fwi2d_amd import configs as C
from cases import overthrust

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fwi2d_amd", "bin")

NT, DT, FMAX = 1001, 1e-3, 30.0
SURVEY = C.Survey("overthrust", 81, 201, 25.0, NT, DT, FMAX, 41, 4, 51, 100.0, 0.0,
                  varVp_rel=0.0, varVp=1000.0, abc_nb=51, abc_factor=0.004, seed=77)
XS = np.array([1000.0, 2000.0, 3000.0, 4000.0], np.float32)
XR = np.linspace(0.0, 5000.0, 51).astype(np.float32)


def run(exe, params):
    print(f"running {exe} {params}", flush=True)
    r = subprocess.run([os.path.join(BIN, exe), params], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-3000:])
    print(r.stderr[-3000:])
    assert r.returncode == 0, (exe, r.returncode)
    return r.stdout


def setup_files(tmp):
    v = overthrust()
    vsmooth = C.smooth_slowness(v, 6.0)
    true_file, init_file = os.path.join(tmp, "true.bin"), os.path.join(tmp, "init.bin")
    v.tofile(true_file)
    vsmooth.tofile(init_file)
    sps, rps = os.path.join(tmp, "sps.txt"), os.path.join(tmp, "rps.txt")
    zs = np.full(len(XS), SURVEY.zsrc, np.float32)
    zr = np.full(len(XR), SURVEY.zrec, np.float32)
    C.write_geometry(sps, rps, XS, zs, XR, zr)
    folders = C.ensure_dirs(tmp)
    prefix = f"seismogram_nt{NT}_nr{len(XR)}_{int(1e6 * DT)}us_shot_"
    params = {}
    for name, model in (("mod", true_file), ("mig", init_file), ("inv", init_file)):
        p = os.path.join(tmp, f"parameters_{name}.txt")
        C.write_parameters(p, SURVEY, model, sps, rps, folders, max_iteration=3,
                           depth_mask=100.0, mig_prefix=prefix, inv_prefix=prefix)
        params[name] = p
    return v, vsmooth, zs, zr, folders, prefix, params


def test_three_mains_end_to_end(tmp_path):
    tmp = str(tmp_path)
    v, vsmooth, zs, zr, folders, prefix, params = setup_files(tmp)
    s = SURVEY
    nrec = len(XR)

    # ---- modeling: one seismogram file per shot, equal to the library path
    run("modeling", params["mod"])
    files = [os.path.join(folders["data"], f"{prefix}{k + 1}.bin") for k in range(len(XS))]
    obs = np.stack([np.fromfile(f, np.float32).reshape(nrec, NT) for f in files])
    w, tlag = F.ricker_wavelet(NT, DT, FMAX)
    eng = F.Engine(s.nz, s.nx, s.dh, NT, DT, abc_nb=s.abc_nb, abc_factor=s.abc_factor,
                   rbc_nb=s.nb, rbc_varVp=s.varVp, rbc_varVp_rel=s.varVp_rel, rbc_seed=s.seed,
                   max_batch=len(XS))
    eng.set_wavelet(w, tlag)
    eng.set_geometry(XS, zs, XR, zr)
    eng.set_model(v)
    ref = eng.modeling(F.ABC)
    assert np.array_equal(obs, ref)
    assert np.all(obs[:, :, 0] == 0.0) and np.abs(obs).max() > 0

    # ---- migration: the RTM section equals the library's post-processed image
    run("migration", params["mig"])
    img_file = os.path.join(folders["seismic"], f"RTM_section_{int(FMAX)}Hz_{s.nz}x{s.nx}_"
                                                f"{int(s.dh)}m.bin")
    img = np.fromfile(img_file, np.float32).reshape(s.nx, s.nz)
    eng.set_model(vsmooth)
    eng.set_observed(obs)
    eng.reset()
    eng.migration()
    ref_img, _, _ = eng.get_image()
    assert np.array_equal(img, ref_img)
    assert np.isfinite(img).all() and np.abs(img).max() > 0

    # ---- inversion: 3 iterations, the reference's outputs, residual goes down
    out = run("inversion", params["inv"])
    suffix = f"{int(FMAX)}Hz_{s.nz}x{s.nx}_{int(s.dh)}m.bin"
    for k in range(1, 4):
        m = np.fromfile(os.path.join(folders["models"], f"model_FWI_iteration_{k}_{suffix}"),
                        np.float32)
        assert m.size == s.nx * s.nz and np.isfinite(m).all()
    final = np.fromfile(os.path.join(folders["models"], f"final_model_FWI_{suffix}"),
                        np.float32).reshape(s.nx, s.nz)
    res = np.loadtxt(os.path.join(folders["residuo"], "convergence_3_iterations.txt"))
    print("residuo", res)
    assert res.shape == (4,)
    assert res[-1] < res[0]
    # the model moved towards the true model
    e0 = np.linalg.norm(vsmooth - v)
    e1 = np.linalg.norm(final - v)
    print(f"model error {e0:.1f} -> {e1:.1f}")
    assert e1 < e0
    assert "residuo" in out
    eng.close()


def test_inversion_resume_equals_straight_run(tmp_path):
    """Three iterations in one run equal two iterations, a checkpoint and a
    resumed third iteration, bit for bit (model and residual history)."""
    tmp = str(tmp_path)
    _, _, _, _, folders, prefix, params = setup_files(tmp)
    run("modeling", params["mod"])
    init_file = os.path.join(tmp, "init.bin")
    sps, rps = os.path.join(tmp, "sps.txt"), os.path.join(tmp, "rps.txt")

    def inv_params(name, iters, extra):
        f = C.ensure_dirs(os.path.join(tmp, name))
        f["data"] = folders["data"]
        p = os.path.join(tmp, f"parameters_{name}_{iters}.txt")
        C.write_parameters(p, SURVEY, init_file, sps, rps, f, max_iteration=iters,
                           depth_mask=100.0, inv_prefix=prefix, extra=extra)
        return p, f

    ck = os.path.join(tmp, "fwi.ckpt")
    pa, fa = inv_params("straight", 3, {"checkpoint_file": os.path.join(tmp, "straight.ckpt")})
    pb1, fb = inv_params("split", 2, {"checkpoint_file": ck})
    pb2, _ = inv_params("split", 3, {"checkpoint_file": ck, "resume": "True"})
    run("inversion", pa)
    run("inversion", pb1)
    out = run("inversion", pb2)
    assert "resumed from" in out
    suffix = f"{int(FMAX)}Hz_{SURVEY.nz}x{SURVEY.nx}_{int(SURVEY.dh)}m.bin"
    a = np.fromfile(os.path.join(fa["models"], f"final_model_FWI_{suffix}"), np.float32)
    b = np.fromfile(os.path.join(fb["models"], f"final_model_FWI_{suffix}"), np.float32)
    assert np.array_equal(a, b)
    ra = open(os.path.join(fa["residuo"], "convergence_3_iterations.txt")).read()
    rb = open(os.path.join(fb["residuo"], "convergence_3_iterations.txt")).read()
    assert ra == rb


def test_prepared_configuration_end_to_end(tmp_path):
    """The data-preparation commands feed the executables: setup of the
    one-shot 121 x 401 configuration, modeling, the direct-wave mute (Python
    and C++ tools, the same bytes),
    migration of the muted data (whose image equals the library's on the
    same muted data), a two-iteration inversion and the viewer's figures."""
    from fwi2d_amd import viewer
    root = str(tmp_path)
    mod, mig, inv = C.main(["setup", "forward-1shot", root, "--max-iteration", "2"])
    s = C.CONFIG1
    run("modeling", mod)
    muted = C.main(["mute", mod])
    assert len(muted) == 1
    # the C++ tool writes the same bytes from the executable's seismogram
    with open(muted[0], "rb") as f:
        py_bytes = f.read()
    r = subprocess.run([os.path.join(BIN, "prep"), "mute", mod], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == muted
    with open(muted[0], "rb") as f:
        assert f.read() == py_bytes
    obs = np.fromfile(muted[0], np.float32).reshape(1, s.nrec, s.nt)
    raw = np.fromfile(muted[0].replace("input_RTM_", ""), np.float32).reshape(s.nrec, s.nt)
    # normalised by the shot's peak, then the direct wave (that peak) tapered
    assert np.isfinite(obs).all() and 0 < np.abs(obs).max() < 1.0
    assert np.allclose(obs[0, :, -1], raw[:, -1] / np.abs(raw).max(), rtol=1e-6, atol=0)

    run("migration", mig)
    suffix = f"{int(s.fmax)}Hz_{s.nz}x{s.nx}_{int(s.dh)}m.bin"
    img = np.fromfile(os.path.join(viewer.read_parameter(mig, "mig_output_folder"),
                                   f"RTM_section_{suffix}"), np.float32).reshape(s.nx, s.nz)
    sps = np.loadtxt(viewer.read_parameter(mig, "SPS"), np.float32, delimiter=",", ndmin=2)
    rps = np.loadtxt(viewer.read_parameter(mig, "RPS"), np.float32, delimiter=",", ndmin=2)
    init = np.fromfile(viewer.read_parameter(mig, "model_file"), np.float32).reshape(s.nx, s.nz)
    w, tlag = F.ricker_wavelet(s.nt, s.dt, s.fmax)
    eng = F.Engine(s.nz, s.nx, s.dh, s.nt, s.dt, abc_nb=s.abc_nb, abc_factor=s.abc_factor,
                   rbc_nb=s.nb, rbc_varVp=s.varVp, rbc_varVp_rel=s.varVp_rel, rbc_seed=s.seed,
                   max_batch=1)
    eng.set_wavelet(w, tlag)
    eng.set_geometry(sps[:, 0], sps[:, 1], rps[:, 0], rps[:, 1])
    eng.set_model(init)
    eng.set_observed(obs)
    eng.reset()
    eng.migration()
    ref_img, _, _ = eng.get_image()
    eng.close()
    assert np.array_equal(img, ref_img)
    assert np.isfinite(img).all() and np.abs(img).max() > 0

    out = run("inversion", inv)
    assert "residuo" in out
    res = np.loadtxt(os.path.join(viewer.read_parameter(inv, "inv_residuo_folder"),
                                  "convergence_2_iterations.txt"))
    assert res.shape == (3,) and np.isfinite(res).all()
    figs = viewer.main([inv, "--out", os.path.join(root, "figures"),
                        "--true", viewer.read_parameter(mod, "model_file")])
    assert sorted(os.path.basename(f) for f in figs) == sorted(
        ["model.png", "seismogram_shot_1.png", "image_RTM.png", "residuo_FWI.png",
         "model_FWI.png"])


def test_autotuned_executables_write_the_same_files(tmp_path):
    """The executables autotune their launch shapes by default
    (gpu_autotune); with gpu_autotune = false they keep the built-in shapes.
    Both write bit-identical seismograms and RTM sections."""
    tmp = str(tmp_path)
    _, _, _, _, folders, prefix, params = setup_files(tmp)
    init_file = os.path.join(tmp, "init.bin")
    true_file = os.path.join(tmp, "true.bin")
    sps, rps = os.path.join(tmp, "sps.txt"), os.path.join(tmp, "rps.txt")
    f2 = C.ensure_dirs(os.path.join(tmp, "fixed"))
    p_mod = os.path.join(tmp, "parameters_mod_fixed.txt")
    C.write_parameters(p_mod, SURVEY, true_file, sps, rps, f2, extra={"gpu_autotune": "false"})
    p_mig = os.path.join(tmp, "parameters_mig_fixed.txt")
    C.write_parameters(p_mig, SURVEY, init_file, sps, rps, f2, mig_prefix=prefix,
                       extra={"gpu_autotune": "false"})
    for p in (params["mod"], p_mod):
        run("modeling", p)
    for p in (params["mig"], p_mig):
        run("migration", p)
    for k in range(1, len(XS) + 1):
        a = np.fromfile(os.path.join(folders["data"], f"{prefix}{k}.bin"), np.float32)
        b = np.fromfile(os.path.join(f2["data"], f"{prefix}{k}.bin"), np.float32)
        assert a.size == len(XR) * NT and np.array_equal(a, b), k
    name = f"RTM_section_{int(FMAX)}Hz_{SURVEY.nz}x{SURVEY.nx}_{int(SURVEY.dh)}m.bin"
    a = np.fromfile(os.path.join(folders["seismic"], name), np.float32)
    b = np.fromfile(os.path.join(f2["seismic"], name), np.float32)
    assert a.size == SURVEY.nz * SURVEY.nx and np.abs(a).max() > 0 and np.array_equal(a, b)


def test_inversion_trial_directions(tmp_path):
    """inv_trial_direction = plain (the control, d = -g) and = network with a
    network of zero weights (d = net(x) - g = -g) run the same iterations and
    write the same files, bit for bit; both log the trial's misfit against
    the plain step's and never keep a trial whose misfit is not below it."""
    tmp = str(tmp_path)
    _, _, _, _, folders, prefix, params = setup_files(tmp)
    run("modeling", params["mod"])
    init_file = os.path.join(tmp, "init.bin")
    sps, rps = os.path.join(tmp, "sps.txt"), os.path.join(tmp, "rps.txt")
    net = F.abi.Net(1, 204, 84, 0)           # 201 x 81 edge-padded to multiples of 4
    zeros = os.path.join(tmp, "zero_network.bin")
    np.zeros(net.nparams, np.float32).tofile(zeros)
    net.close()
    outs = {}
    for name, extra in (("plain", {"inv_trial_direction": "plain"}),
                        ("network", {"inv_trial_direction": "network",
                                     "inv_trial_network": zeros,
                                     "inv_trial_multiples": "2,4"})):
        f = C.ensure_dirs(os.path.join(tmp, name))
        f["data"] = folders["data"]
        p = os.path.join(tmp, f"parameters_{name}.txt")
        C.write_parameters(p, SURVEY, init_file, sps, rps, f, max_iteration=3,
                           depth_mask=100.0, inv_prefix=prefix, extra=extra)
        out = run("inversion", p)
        trials = [l for l in out.splitlines() if l.startswith("trial direction")]
        assert len(trials) >= 1
        for line in trials:
            words = line.split()
            fl, fp = float(words[3]), float(words[9])
            assert ("trial taken" in line) == (fl < fp), line
        suffix = f"{int(FMAX)}Hz_{SURVEY.nz}x{SURVEY.nx}_{int(SURVEY.dh)}m.bin"
        outs[name] = (np.fromfile(os.path.join(f["models"], f"final_model_FWI_{suffix}"),
                                  np.float32),
                      open(os.path.join(f["residuo"], "convergence_3_iterations.txt")).read())
    assert np.array_equal(outs["plain"][0], outs["network"][0])
    assert outs["plain"][1] == outs["network"][1]


def test_reference_random_boundary_keys(tmp_path):
    """A parameters.txt with the reference's keys only (no rbc_kind): the
    executables take the reference's boundary, Gaussian noise of standard
    deviation mig_rbc_varVp plus Poisson-disk blobs of radius up to
    mig_rbc_ratio (seeded), and the RTM section equals the library's with
    that boundary bit for bit."""
    tmp = str(tmp_path)
    v, vsmooth, zs, zr, folders, prefix, params = setup_files(tmp)
    s = SURVEY
    for name in ("mod", "mig"):
        with open(params[name]) as f:
            lines = [l for l in f if not l.startswith("rbc_kind")]
        with open(params[name], "w") as f:
            f.writelines(lines)
    run("modeling", params["mod"])
    nrec = len(XR)
    files = [os.path.join(folders["data"], f"{prefix}{k + 1}.bin") for k in range(len(XS))]
    obs = np.stack([np.fromfile(f, np.float32).reshape(nrec, NT) for f in files])
    run("migration", params["mig"])
    img = np.fromfile(os.path.join(folders["seismic"], f"RTM_section_{int(FMAX)}Hz_{s.nz}x"
                                   f"{s.nx}_{int(s.dh)}m.bin"), np.float32).reshape(s.nx, s.nz)
    w, tlag = F.ricker_wavelet(NT, DT, FMAX)
    eng = F.Engine(s.nz, s.nx, s.dh, NT, DT, abc_nb=s.abc_nb, abc_factor=s.abc_factor,
                   rbc_nb=s.nb, rbc_varVp=s.varVp, rbc_varVp_rel=0.0, rbc_seed=s.seed,
                   max_batch=len(XS), rbc_kind=F.abi.RBC_GAUSSIAN, rbc_ratio=200.0)
    eng.set_wavelet(w, tlag)
    eng.set_geometry(XS, zs, XR, zr)
    eng.set_model(vsmooth)
    eng.set_observed(obs)
    eng.reset()
    eng.migration()
    ref_img, _, _ = eng.get_image()
    eng.close()
    assert np.array_equal(img, ref_img)
    assert np.isfinite(img).all() and np.abs(img).max() > 0
