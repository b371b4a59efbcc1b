#!/usr/bin/env python3
"""Benchmark: batched reverse-time migration, 64 shots, on one GPU.

Workload (one "step" = one full RTM of the 64-shot batch, inputs resident in
HBM): grid 301 x 1001 at 10 m with a 150-cell random-boundary pad
(601 x 1301 cells), nt = 4001 at dt = 0.8 ms (+ tlag), Ricker 10 Hz peak
(max_frequency 30 Hz), 64 shots along the surface, 1001 receivers at 10 m,
six seeded dipping layers 1500-4000 m/s, migration velocity = the true model
smoothed (sigma 20 cells, slowness).  Observed data are modeled from the true
model (sponge boundary) and direct-wave muted before the timed region.

Per step: random-boundary generation, one forward pass and one reverse pass
(source reversal + receiver back-propagation + imaging) over all 64 shots.
Cell-updates = shots x nzz x nxx x T per pass, 2 per cell in the reverse pass.

Prints one JSON line (the metric is Gcell-updates/s).  `roofline` is the
reverse kernel (the dominant one), `roofline_forward` the forward kernel.
roofline.traffic is the HBM bytes per launch from two rocprofv3 --pmc passes
(FETCH_SIZE, then WRITE_SIZE) over a short run of the same kernels in a child
process, each scaled by what the counter reads for a copy of known size;
roofline.achieved / frac are those bytes over the launch time (HIP events on
the engine's stream) against the 8 TB/s peak.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_FWD = 16.0      # read u_t, u_{t-1}, v; write u_{t+1}  (FP32)
BYTES_REV = 22.0      # per cell-update of the reverse pass (44 B per cell-step)
# non-packed FP32 VALU issue: one wave64 instruction per cycle per CU
# (4 SIMD16 units), 256 CUs at the 2.4 GHz peak engine clock
VALU_PEAK = 256 * 2.4e9       # wave-instructions / s
PMC_NT = 300                  # time samples of the counter-collection run
PMC_COPY_BYTES = 256 << 20    # calibration copy (reads and writes this many bytes)
PMC_TIMEOUT = 240             # seconds per counter pass


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--shots", type=int, default=64)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--cpu-shots", type=int, default=8)
    p.add_argument("--nt", type=int, default=0,
                   help="override time_samples (short profiling runs; same per-step work)")
    p.add_argument("--no-autotune", action="store_true")
    p.add_argument("--fwd-tuning", default="", help="steps,vz,xw of the forward kernel")
    p.add_argument("--rev-tuning", default="", help="steps,vz,xw of the reverse kernel")
    p.add_argument("--graph", type=int, default=-1)
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the rocprofv3 counter passes (roofline.traffic = null)")
    p.add_argument("--workload", default="rtm64",
                   choices=("rtm64", "forward1", "rtm1", "fwi128", "grad256", "net"),
                   help="rtm64 is the headline; forward1, fwi128 and grad256 show the rate's "
                        "dependence on shape; net times the learned preconditioner's Adam step")
    p.add_argument("--pmc-child", action="store_true",
                   help="internal: the short run profiled by the counter passes")
    return p.parse_args()


def cpu_baseline(s, vmig, xs, zs, xr, zr, w, tlag, obs, threads, nshots):
    """FP32 CPU statement of the reference stencil (OpenMP) on the first
    `nshots` shots of the same workload: random-boundary forward + reverse
    imaging pass per shot."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    nb = s.nb
    vpad = ob.expand(vmig, nb)
    gi = lambda x: ob.grid_index(float(x), np.float32(s.dh), nb)
    riz = np.array([gi(z) for z in zr], np.int32)
    rix = np.array([gi(x) for x in xr], np.int32)
    one = np.float32(1.0)
    rel = np.float32(s.varVp_rel)
    vlo = vmig.min() * (one - rel) - np.float32(s.varVp)
    vhi = vmig.max() * (one + rel) + np.float32(s.varVp)
    vcfl = np.float32(np.float32(0.999) * np.float32(0.5546) * np.float32(s.dh)) / np.float32(s.dt)
    vhi = min(vhi, vcfl)
    T = s.nt + tlag
    nzz, nxx = s.nz + 2 * nb, s.nx + 2 * nb
    acc = np.zeros((nxx, nzz), np.float32)
    ill = np.zeros((nxx, nzz), np.float32)
    el = 0.0
    for k in range(nshots):
        vrb = ob.rb_velocity(vpad, nb, np.float32(s.varVp), rel, vlo, vhi, s.seed, k)
        t0 = time.perf_counter()
        _, P, Q = ob.forward(vrb, s.dh, s.dt, s.nt, tlag, w, gi(zs[k]), gi(xs[k]), riz, rix,
                             precision="f32", nthreads=threads)
        ob.reverse(vrb, s.dh, s.dt, s.nt, tlag, P, Q, riz, rix, obs[k], 0, acc=acc, illum=ill,
                   precision="f32", nthreads=threads)
        el += time.perf_counter() - t0
    cells = nshots * nzz * nxx * T * 3.0
    return {"value": round(cells / el / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{nshots} of {s.nsrc} shots, full RTM each (RB forward + reverse "
                      f"imaging), {nzz}x{nxx} cells x {T} steps, FP32 OpenMP, "
                      f"{el:.1f} s of CPU time",
            "host": host_cpu()}


def host_cpu():
    """CPU model and logical CPU count of this host (Linux), for the record."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs"



def compulsory_bpc(kind, steps):
    """Bytes per cell one launch must move however it is tiled: reverse
    (kind 1) reads S, S_old, R, R_old, c, image, illum and writes S, R (two
    fields each when steps >= 2), image, illum; forward reads u, u_old, c and
    writes u (two fields when steps >= 2)."""
    if kind == 1:
        return 52.0 if steps >= 2 else 44.0
    return 20.0 if steps >= 2 else 16.0


def roof(ms, n, cells, bpc, copy_gbs, kind=0, steps=1):
    """Roofline record of one kind of step launch from the engine's event
    profile: ms over n launches that advanced `cells` cell-updates.

    `achieved` / `frac` are HBM bytes per launch over the launch time against
    the 8 TB/s peak.  Until the counter passes have run (set_traffic), the
    bytes are the launch's compulsory bytes: what a launch of `steps` steps
    must move however it is tiled (each input field read once, each output
    written once), a lower bound on its traffic.  `per_step_equivalent` is the
    algorithmic figure of unblocked steps (bpc bytes per cell-update times the
    cell-updates of the launch): a launch that advances several steps stands
    for that many bytes without moving them, so its `frac` can pass 1 and is
    not a fraction of the HBM roofline."""
    avg_s = ms / n * 1e-3
    per_launch = cells / n
    cells_launch = per_launch / steps / (2.0 if kind == 1 else 1.0)
    comp = cells_launch * compulsory_bpc(kind, steps)
    alg_gbs = per_launch * bpc / avg_s / 1e9
    comp_gbs = comp / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(comp_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(comp_gbs / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_basis": "compulsory bytes per launch (counters not run)",
            "kernel_avg_us": round(avg_s * 1e6, 2),
            "gcell_per_s": round(per_launch / avg_s / 1e9, 2),
            "frac_of_measured_copy": round(comp_gbs / copy_gbs, 4),
            "measured_copy_GBs": round(copy_gbs, 1),
            "steps_per_launch": steps,
            "cell_updates_per_launch": per_launch,
            "compulsory_bytes_per_launch": comp,
            "compulsory_frac": round(comp_gbs / HBM_PEAK_GBS, 4),
            "per_step_equivalent": {
                "algorithmic_bytes_per_launch": per_launch * bpc,
                "bytes_per_cell_update": bpc,
                "achieved": round(alg_gbs, 1), "frac": round(alg_gbs / HBM_PEAK_GBS, 4)}}


def set_traffic(r, read, write):
    """Switches a roofline record to the counter-measured HBM bytes per launch
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated): `achieved` and `frac` become
    the kernel's physical HBM rate and its fraction of the peak."""
    total = read + write
    avg_s = r["kernel_avg_us"] * 1e-6
    gbs = total / avg_s / 1e9
    r["traffic"] = round(total, 0)
    r["traffic_read"] = round(read, 0)
    r["traffic_write"] = round(write, 0)
    r["achieved"] = round(gbs, 1)
    r["frac"] = round(gbs / HBM_PEAK_GBS, 4)
    r["frac_of_measured_copy"] = round(gbs / r["measured_copy_GBs"], 4)
    r["bytes_basis"] = "HBM traffic per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE)"
    r["traffic_over_compulsory"] = round(total / r["compulsory_bytes_per_launch"], 4)


def other_workload(a):
    """The other benchmark configurations, for the sensitivity of the rate to
    the problem shape (same JSON keys; not the headline):
      forward1  1 shot, 121x401 + 100-cell random boundary, nt 2001, 1 ms,
                15 Hz peak, 401 receivers (the field fits in the caches)
      rtm1      the RTM of that shot (observed data modeled with a 100-cell
                sponge on the true model, migrated on its smoothed-slowness
                starting model)
      fwi128    one plain FWI iteration (gradient pass + two line-search
                misfits), 128 shots, 301x1001 + 150 (random) / 100 (sponge),
                nt 4001, 0.8 ms, 8 Hz peak, 1001 receivers
      grad256   one FWI gradient pass, 256 shots, 401x1601 + 200 (random) /
                100 (sponge), nt 5001, 0.6 ms, 12 Hz peak, 1601 receivers"""
    import dataclasses
    import fwi2d_amd as F
    from fwi2d_amd import configs as C
    copy_gbs = F.abi.stream_copy(1 << 30, reps=20)
    if a.workload in ("forward1", "rtm1"):
        s = C.CONFIG1
        true_v = C.config1_model()
        init_v = None if a.workload == "forward1" else C.smooth_slowness(true_v, 3.0)
        s = dataclasses.replace(s, seed=2024)
    elif a.workload == "fwi128":
        s = C.CONFIG3
        true_v, init_v = C.config3_models()
    else:
        s = C.CONFIG5
        true_v, init_v = C.config5_models(5, s.nz, s.nx)
    xs, zs, xr, zr = s.acquisition()
    w, tlag = F.ricker_wavelet(s.nt, s.dt, s.fmax)
    T = s.nt + tlag
    eng = F.Engine(s.nz, s.nx, s.dh, s.nt, s.dt, abc_nb=0 if init_v is None else 100,
                   abc_factor=0.0 if init_v is None else 0.0035, rbc_nb=s.nb,
                   rbc_varVp=s.varVp, rbc_varVp_rel=s.varVp_rel, rbc_seed=s.seed,
                   max_batch=s.nsrc)
    eng.set_wavelet(w, tlag)
    eng.set_geometry(xs, zs, xr, zr)
    eng.set_model(true_v)
    nzz, nxx, _ = eng.grid_shape(F.RBC)
    if init_v is not None:
        eng.set_observed(eng.modeling(F.ABC))
        eng.set_model(init_v)
        azz, axx, _ = eng.grid_shape(F.ABC)
    if a.fwd_tuning or a.rev_tuning:
        # pinned shapes (profiling runs); the forward one also for the sponge kernel
        for kernels, spec in (((0, 2), a.fwd_tuning), ((1,), a.rev_tuning)):
            if spec:
                st, vz, xw = (int(x) for x in spec.split(","))
                for kernel in kernels:
                    eng.set_tuning(vz, xw, kernel=kernel, steps=st)
    elif not a.no_autotune:
        eng.autotune(s.nsrc, reps=8)

    def step():
        if a.workload == "forward1":
            eng.modeling(F.RBC)
            return s.nsrc * nzz * nxx * T
        if a.workload == "rtm1":
            eng.reset(F.RESET_IMAGE | F.RESET_ILLUM)
            eng.migration()
            return s.nsrc * nzz * nxx * T * 3.0
        eng.reset(F.RESET_GRADIENT)
        eng.gradient(iteration=1)
        cells = s.nsrc * (azz * axx + nzz * nxx * 3.0) * T
        if a.workload == "fwi128":
            for scale in (0.99, 0.98):               # the two line-search trial models
                eng.set_model((init_v * scale).astype(np.float32))
                eng.misfit()
            eng.set_model(init_v)
            cells += 2.0 * s.nsrc * azz * axx * T
        return cells

    for _ in range(a.warmup):
        step()
    eng.synchronize()
    eng.profile_enable(True)
    t0 = time.perf_counter()
    cells = sum(step() for _ in range(a.steps))
    eng.synchronize()
    el = time.perf_counter() - t0
    fwd = eng.profile_read(0)
    rev = eng.profile_read(1)
    eng.profile_enable(False)
    tune = {name: dict(zip(("steps_per_launch", "vz", "xw"), eng.get_tuning(k)[:3]))
            for k, name in ((0, "forward"), (1, "reverse"), (2, "forward_sponge"))}
    eng.close()
    out = {"metric": "Gcell-updates/s", "value": round(cells / el / 1e9, 2),
           "unit": "Gcell-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": round(el / a.steps * 1e3, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": a.workload, "shots": s.nsrc, "nzz": nzz, "nxx": nxx,
                      "steps_per_pass": T, "tuning": tune,
                      "measured_copy_GBs": round(copy_gbs, 1)},
           # forward launches of both boundaries share one profile slot
           "roofline_forward": roof(*fwd, BYTES_FWD, copy_gbs, 0,
                                    tune["forward"]["steps_per_launch"])}
    if rev[1]:
        out["roofline"] = roof(*rev, BYTES_REV, copy_gbs, 1, tune["reverse"]["steps_per_launch"])
    print(json.dumps(out), flush=True)


def net_workload(a):
    """The learned preconditioner's network on its training shape: batch 16
    of 3 x 384 x 128 inputs (seeded random), `steps` Adam steps (forward,
    backward, weight update) after `warmup`, then as many forward passes."""
    from fwi2d_amd import abi as A
    B, H, W = 16, 384, 128
    rng = np.random.default_rng(0)
    x = rng.standard_normal((B, 3, H, W)).astype(np.float32)
    t = (0.1 * rng.standard_normal((B, H, W))).astype(np.float32)
    net = A.Net(B, H, W, 0)
    for _ in range(a.warmup):
        net.train_step(x, t, 1e-3)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        net.train_step(x, t, 1e-3)
    el = (time.perf_counter() - t0) / a.steps
    t0 = time.perf_counter()
    for _ in range(a.steps):
        net.forward(x)
    fw = (time.perf_counter() - t0) / a.steps
    print(json.dumps({"metric": "preconditioner Adam steps/s", "value": round(1.0 / el, 2),
                      "unit": "steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
                      "ms_per_step": round(el * 1e3, 3), "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
                      "data": "synthetic (seeded random inputs and targets)",
                      "config": {"workload": "net", "batch": B, "channels_in": 3, "height": H,
                                 "width": W, "forward_ms": round(fw * 1e3, 3),
                                 "note": "host copies of the batch included"}}), flush=True)


def kernel_key(kind, steps, vz):
    """Name fragment of the random-boundary forward (kind 0) or imaging
    reverse (kind 1) kernel instance (either staging build); vz 16 is the
    forward kernel on 64 x 64-cell LDS tiles."""
    if kind == 0:
        if vz == 16:
            return f"fwdt_kernel<{steps}, false,"
        return f"fwd{steps}_kernel<{vz}, false,"
    return f"rev{steps}_kernel<{vz}, 0,"


def counter_pass(prof, counter, child_args, workdir):
    """One rocprofv3 --pmc pass over the child run; returns {kernel: [KiB per
    dispatch]} or raises RuntimeError."""
    import csv
    import glob
    import signal
    import subprocess
    out = os.path.join(workdir, counter.lower())
    cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "pmc", "--",
           sys.executable, os.path.abspath(__file__), "--pmc-child"] + child_args
    log = open(os.path.join(workdir, counter.lower() + ".log"), "w")
    p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=log, start_new_session=True,
                         text=True)
    try:
        so, _ = p.communicate(timeout=PMC_TIMEOUT)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.communicate()
        raise RuntimeError(f"{counter} pass timed out")
    finally:
        log.close()
    # rocprofv3 may exit non-zero after a complete run; the child's own line
    # and the counter file decide
    if "pmc-child done" not in so:
        raise RuntimeError(f"{counter} pass: child did not finish (rc {p.returncode})")
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"{counter} pass: no counter file")
    vals = {}
    with open(files[0], newline="") as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name") != counter:
                continue
            vals.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    return vals


def measure_traffic(nshots, tuning):
    """HBM bytes per launch of the forward and reverse kernels (FETCH_SIZE and
    WRITE_SIZE, each times the copy-calibrated factor).  Returns (dict, note)."""
    import shutil
    import tempfile
    prof = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3"
                                         if os.path.exists("/opt/rocm/bin/rocprofv3") else None)
    if prof is None:
        return None, "rocprofv3 not found"
    (stf, vzf, xwf), (str_, vzr, xwr) = tuning
    child = ["--nt", str(PMC_NT), "--shots", str(nshots), "--graph", "0",
             "--fwd-tuning", f"{stf},{vzf},{xwf}", "--rev-tuning", f"{str_},{vzr},{xwr}"]
    want = {"forward": kernel_key(0, stf, vzf), "reverse": kernel_key(1, str_, vzr)}
    res = {"forward": {}, "reverse": {}, "calibration": {}}
    valu = {}
    with tempfile.TemporaryDirectory(prefix="fwi_pmc_") as wd:
        try:       # VALU instructions issued per dispatch (the compute side)
            vals = counter_pass(prof, "SQ_INSTS_VALU", child, wd)
            for kind, pre in want.items():
                vs = [v for k, v_ in vals.items() if pre in k for v in v_]
                if vs:
                    valu[kind] = float(np.median(vs))
        except RuntimeError:
            valu = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            vals = counter_pass(prof, counter, child, wd)
            copy = [v for k, vs in vals.items() if "copy_kernel" in k for v in vs]
            if not copy:
                raise RuntimeError(f"{counter}: copy kernel not seen")
            factor = (PMC_COPY_BYTES / 1024.0) / float(np.median(copy))
            res["calibration"][counter] = round(factor, 4)
            for kind, pre in want.items():
                vs = [v for k, v_ in vals.items() if pre in k for v in v_]
                if not vs:
                    raise RuntimeError(f"{counter}: {pre} not seen")
                res[kind][counter] = float(np.median(vs)) * factor * 1024.0
    out = {k: res[k]["FETCH_SIZE"] + res[k]["WRITE_SIZE"] for k in ("forward", "reverse")}
    note = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, median KiB per dispatch over a "
            f"{PMC_NT}-sample run of the same kernels and launch shapes, scaled by the "
            f"counter's reading for a {PMC_COPY_BYTES >> 20} MiB copy (factors "
            f"{res['calibration']})")
    return {"bytes": out, "read": {k: res[k]["FETCH_SIZE"] for k in out},
            "write": {k: res[k]["WRITE_SIZE"] for k in out}, "valu": valu}, note


def pmc_child(a):
    """The run profiled by the counter passes: a calibration copy of known
    size, then one short migration with the given launch shapes."""
    import fwi2d_amd as F
    from fwi2d_amd import configs as C
    import dataclasses
    s = dataclasses.replace(C.CONFIG2, nt=a.nt)
    true_v, mig_v = C.config2_models()
    xs, zs, xr, zr = s.acquisition()
    if a.shots != s.nsrc:
        sel = np.linspace(0, s.nsrc - 1, a.shots).round().astype(int)
        xs, zs = xs[sel], zs[sel]
    w, tlag = F.ricker_wavelet(s.nt, s.dt, s.fmax)
    F.abi.stream_copy(PMC_COPY_BYTES, reps=3)
    eng = F.Engine(s.nz, s.nx, s.dh, s.nt, s.dt, abc_nb=s.abc_nb, abc_factor=s.abc_factor,
                   rbc_nb=s.nb, rbc_varVp=s.varVp, rbc_varVp_rel=s.varVp_rel, rbc_seed=s.seed,
                   max_batch=a.shots)
    eng.set_wavelet(w, tlag)
    eng.set_geometry(xs, zs, xr, zr)
    eng.set_model(true_v)
    obs = eng.modeling(F.ABC)
    eng.set_observed(obs)
    eng.set_model(mig_v)
    for kernel, spec in ((0, a.fwd_tuning), (1, a.rev_tuning)):
        st, vz, xw = (int(x) for x in spec.split(","))
        eng.set_tuning(vz, xw, 0, kernel=kernel, steps=st)
    eng.migration()
    eng.synchronize()
    eng.close()
    print("pmc-child done", flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the shot axis "
              f"would shard across GPUs, but this build measures one GPU's share.",
              file=sys.stderr)
        sys.exit(2)
    if a.pmc_child:
        return pmc_child(a)
    if a.workload == "net":
        return net_workload(a)
    if a.workload != "rtm64":
        return other_workload(a)
    import fwi2d_amd as F
    from fwi2d_amd import configs as C

    s = C.CONFIG2
    if a.nt:
        import dataclasses
        s = dataclasses.replace(s, nt=a.nt)
    nshots = a.shots
    true_v, mig_v = C.config2_models()
    xs, zs, xr, zr = s.acquisition()
    if nshots != s.nsrc:
        sel = np.linspace(0, s.nsrc - 1, nshots).round().astype(int)
        xs, zs = xs[sel], zs[sel]
    w, tlag = F.ricker_wavelet(s.nt, s.dt, s.fmax)
    T = s.nt + tlag

    copy_gbs = F.abi.stream_copy(1 << 30, reps=20)    # practical HBM ceiling (1 GiB copy)
    eng = F.Engine(s.nz, s.nx, s.dh, s.nt, s.dt, abc_nb=s.abc_nb, abc_factor=s.abc_factor,
                   rbc_nb=s.nb, rbc_varVp=s.varVp, rbc_varVp_rel=s.varVp_rel, rbc_seed=s.seed,
                   max_batch=nshots)
    eng.set_tuning(use_graph=a.graph)
    eng.set_wavelet(w, tlag)
    eng.set_geometry(xs, zs, xr, zr)
    # observed data: true model, sponge boundary, direct-wave mute (setup, untimed)
    eng.set_model(true_v)
    obs = eng.modeling(F.ABC)
    vtop = float(true_v[:, 0].min())
    for k in range(nshots):
        obs[k] = C.mute_direct_wave(obs[k], xs[k], zs[k], xr, zr, s.dt, vtop,
                                    delay=tlag * s.dt + 0.05)
    t_up = time.perf_counter()
    eng.set_observed(obs)                  # synchronous host-to-device copy
    t_up = time.perf_counter() - t_up
    eng.set_model(mig_v)
    nzz, nxx, _ = eng.grid_shape(F.RBC)
    if a.fwd_tuning or a.rev_tuning:
        for kernel, spec in ((0, a.fwd_tuning), (1, a.rev_tuning)):
            if spec:
                st, vz, xw = (int(x) for x in spec.split(","))
                eng.set_tuning(vz, xw, a.graph, kernel=kernel, steps=st)
    elif not a.no_autotune:
        eng.autotune(nshots, reps=8)

    for _ in range(a.warmup):
        eng.reset(F.RESET_IMAGE | F.RESET_ILLUM)
        eng.migration()
    eng.synchronize()
    eng.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        eng.reset(F.RESET_IMAGE | F.RESET_ILLUM)
        eng.migration()
    eng.synchronize()
    el = time.perf_counter() - t0
    fwd_ms, fwd_n, fwd_cells = eng.profile_read(0)
    rev_ms, rev_n, rev_cells = eng.profile_read(1)
    eng.profile_enable(False)
    t_dn = time.perf_counter()
    img, _, _ = eng.get_image()
    t_dn = time.perf_counter() - t_dn
    if not np.all(np.isfinite(img)):
        raise RuntimeError("non-finite image")

    cells_step = nshots * nzz * nxx * T * 3.0
    value = cells_step * a.steps / el / 1e9
    # one RTM with its host boundary: the observed data uploaded from host
    # memory and the image brought back (reported beside `value`, never as it)
    pcie = {"value": round(cells_step / (el / a.steps + t_up + t_dn) / 1e9, 2),
            "unit": "Gcell-updates/s", "upload_s": round(t_up, 4),
            "upload_bytes": int(obs.nbytes), "download_s": round(t_dn, 4),
            "note": "one RTM including the host-to-device copy of the observed data "
                    "(pageable host memory) and the image download with its "
                    "post-processing"}

    stf, vzf, xwf, g = eng.get_tuning(0)
    str_, vzr, xwr, _ = eng.get_tuning(1)
    sts, vzs, xws, _ = eng.get_tuning(2)
    rev_roof = roof(rev_ms, rev_n, rev_cells, BYTES_REV, copy_gbs, 1, str_)
    fwd_roof = roof(fwd_ms, fwd_n, fwd_cells, BYTES_FWD, copy_gbs, 0, stf)
    rev_roof["kernel"] = f"rev{str_}_kernel<{vzr}, 0, staged> (fused reverse + imaging)"
    fwd_roof["kernel"] = f"fwd{stf}_kernel<{vzf}, false, staged> (random-boundary forward)"
    out = {
        "metric": "Gcell-updates/s", "value": round(value, 2), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(el / a.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded dipping-layer model, modeled + muted observed data)",
        "config": {"workload": "batched RTM, 64 shots, 301x1001 grid + 150-cell random "
                               "boundary (601x1301), nt=4001+%d, dt=0.8 ms" % tlag,
                   "shots": nshots, "nzz": nzz, "nxx": nxx, "steps_per_pass": T,
                   "tuning": {"forward": {"steps_per_launch": stf, "vz": vzf, "xw": xwf},
                              "reverse": {"steps_per_launch": str_, "vz": vzr, "xw": xwr},
                              "forward_sponge": {"steps_per_launch": sts, "vz": vzs, "xw": xws},
                              "graph": g},
                   "measured_copy_GBs": round(copy_gbs, 1)},
        "roofline": rev_roof,
        "roofline_forward": fwd_roof,
        "pcie_inclusive": pcie,
    }
    eng.close()
    if not a.no_traffic:
        try:
            tr, note = measure_traffic(nshots, ((stf, vzf, xwf), (str_, vzr, xwr)))
        except RuntimeError as e:
            tr, note = None, f"counter passes failed: {e}"
        if tr is not None:
            for key, kind in (("roofline", "reverse"), ("roofline_forward", "forward")):
                set_traffic(out[key], tr["read"][kind], tr["write"][kind])
                if kind in tr["valu"]:
                    n_valu = tr["valu"][kind]
                    rate = n_valu / (out[key]["kernel_avg_us"] * 1e-6)
                    out[key]["valu"] = {"insts_per_launch": n_valu, "achieved": round(rate, 0),
                                        "peak": VALU_PEAK, "unit": "wave-instructions/s",
                                        "frac": round(rate / VALU_PEAK, 4)}
        out["roofline"]["traffic_method"] = note
    if not a.no_cpu_baseline:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(s, mig_v, xs, zs, xr, zr, w, tlag, obs, threads,
                                           min(a.cpu_shots, nshots))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
