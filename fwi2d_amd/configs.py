"""Seeded survey generators and data preparation for the benchmark configs.

Restates the reference's Python harness (run/generate_setup.py:23-53 --
acquisition files and the smoothed-slowness starting model;
run/generate_input_data.py:22-46 -- direct-wave mute) and adds the seeded
model families the benchmark configurations name (flat / dipping layers,
Gaussian lenses, salt-like polygons).  Every model is [nx, nz] float32 with z
fastest, the reference's file layout.

The command line (`python -m fwi2d_amd.configs setup | init | mute`, see
`main`) writes the files the three executables read: a named
configuration's models, geometry and parameter files, a smoothed starting
model of any model file, and the muted migration inputs.
"""
import dataclasses
import os

import numpy as np


@dataclasses.dataclass
class Survey:
    name: str
    nz: int
    nx: int
    dh: float
    nt: int
    dt: float
    fmax: float                   # the reference's max_frequency = 3 x Ricker peak
    nb: int                       # pad cells per side (4-sided)
    nsrc: int
    nrec: int
    zsrc: float
    zrec: float
    varVp_rel: float = 0.4        # random boundary: +-40 % ramp
    varVp: float = 0.0
    abc_nb: int = 100
    abc_factor: float = 0.0035
    seed: int = 2024

    def acquisition(self):
        """Sources evenly spaced along the surface, one receiver per column
        band (run/generate_setup.py:31-36 spaces them with linspace)."""
        xmax = (self.nx - 1) * self.dh
        if self.nsrc == 1:
            xs = np.array([xmax / 2.0])
        else:
            xs = np.linspace(0.02 * xmax, 0.98 * xmax, self.nsrc)
        xr = np.linspace(0.0, xmax, self.nrec)
        return (xs.astype(np.float32), np.full(self.nsrc, self.zsrc, np.float32),
                xr.astype(np.float32), np.full(self.nrec, self.zrec, np.float32))


# The five benchmark configurations (sizes as stated; peaks -> fmax = 3 x peak).
CONFIG1 = Survey("forward-1shot", 121, 401, 10.0, 2001, 1e-3, 45.0, 100, 1, 401, 20.0, 10.0)
CONFIG2 = Survey("rtm-64shot", 301, 1001, 10.0, 4001, 8e-4, 30.0, 150, 64, 1001, 20.0, 10.0)
CONFIG3 = Survey("fwi-128shot", 301, 1001, 10.0, 4001, 8e-4, 24.0, 150, 128, 1001, 20.0, 10.0)
CONFIG5 = Survey("fwi-256shot", 401, 1601, 10.0, 5001, 6e-4, 36.0, 200, 256, 1601, 20.0, 10.0)


def gaussian_smooth(a, sigma, truncate=4.0):
    """Separable Gaussian filter with reflected edges, the semantics of
    scipy.ndimage.gaussian_filter(mode='reflect', truncate=4.0) that
    run/generate_setup.py:49 uses."""
    a = np.asarray(a, np.float64)
    r = int(truncate * sigma + 0.5)
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    k /= k.sum()
    out = a
    for axis in range(a.ndim):
        pad = [(0, 0)] * a.ndim
        pad[axis] = (r, r)
        p = np.pad(out, pad, mode="symmetric")
        acc = np.zeros_like(out)
        for i, w in enumerate(k):
            sl = [slice(None)] * a.ndim
            sl[axis] = slice(i, i + out.shape[axis])
            acc += w * p[tuple(sl)]
        out = acc
    return out


def smooth_slowness(v, sigma, keep_rows=0):
    """1 / gaussian(1/v, sigma) with the top `keep_rows` rows copied from v
    (run/generate_setup.py:49-51)."""
    out = 1.0 / gaussian_smooth(1.0 / np.asarray(v, np.float64), sigma)
    if keep_rows > 0:
        out[:, :keep_rows] = v[:, :keep_rows]
    return out.astype(np.float32)


def flat_layers(nz, nx, speeds):
    v = np.zeros((nx, nz), np.float32)
    n = len(speeds)
    for k, s in enumerate(speeds):
        v[:, k * nz // n:] = s
    return v


def gaussian_lens(v, zc, xc, sigma, amp):
    nx, nz = v.shape
    z = np.arange(nz)[None, :]
    x = np.arange(nx)[:, None]
    v += (amp * np.exp(-0.5 * ((z - zc) ** 2 + (x - xc) ** 2) / sigma ** 2)).astype(np.float32)
    return v


def dipping_layers(nz, nx, nlayers, vmin, vmax, seed):
    """nlayers seeded dipping interfaces, speeds increasing with depth."""
    rng = np.random.default_rng(seed)
    speeds = np.linspace(vmin, vmax, nlayers)
    base = np.sort(rng.uniform(0.1, 0.95, nlayers - 1)) * nz
    dips = rng.uniform(-0.12, 0.12, nlayers - 1)
    x = np.arange(nx)
    v = np.full((nx, nz), speeds[0], np.float32)
    z = np.arange(nz)[None, :]
    for k in range(nlayers - 1):
        depth = base[k] + dips[k] * (x - nx / 2.0)
        v[z >= depth[:, None]] = speeds[k + 1]
    return v


def salt_polygon(v, zc, xc, hz, hx, speed, seed, nvert=12):
    """Seeded random star-shaped polygon of half-sizes (hz, hx) cells."""
    rng = np.random.default_rng(seed)
    ang = np.sort(rng.uniform(0, 2 * np.pi, nvert))
    rad = rng.uniform(0.6, 1.0, nvert)
    nx, nz = v.shape
    z = np.arange(nz)[None, :] - zc
    x = np.arange(nx)[:, None] - xc
    theta = np.arctan2(z / hz, x / hx) % (2 * np.pi)
    r = np.sqrt((z / hz) ** 2 + (x / hx) ** 2)
    # boundary radius by linear interpolation between vertices
    a2 = np.concatenate([ang - 2 * np.pi, ang, ang + 2 * np.pi])
    r2 = np.concatenate([rad, rad, rad])
    rb = np.interp(theta, a2, r2)
    v[r <= rb] = speed
    return v


def config1_model(seed=1):
    v = flat_layers(121, 401, [1500, 2000, 2500, 3000])
    rng = np.random.default_rng(seed)
    return gaussian_lens(v, rng.uniform(40, 90), rng.uniform(120, 280), 15.0, 250.0)


def config2_models(seed=2):
    true = dipping_layers(301, 1001, 6, 1500.0, 4000.0, seed)
    return true, smooth_slowness(true, 20.0)


def config3_models(seed=3):
    rng = np.random.default_rng(seed)
    v = flat_layers(301, 1001, list(np.linspace(1500, 3500, 6)))
    for k in range(3):
        gaussian_lens(v, rng.uniform(80, 250), rng.uniform(100, 900), 25.0,
                      300.0 if k % 2 == 0 else -300.0)
    salt_polygon(v, rng.uniform(150, 220), rng.uniform(350, 650), 30, 100, 4500.0, seed + 7)
    return v, smooth_slowness(v, 30.0)


def config4_models(seed, nz=256, nx=768, sigma=None):
    """Training family (config-2 style): 4-7 seeded dipping layers from 1500 m/s
    to 3500-4500 m/s, 0-3 Gaussian lenses (+-300 m/s) and, every other seed, a
    salt-like body; starting model = smoothed slowness (sigma defaults to 20
    cells at nz = 301, scaled with nz)."""
    rng = np.random.default_rng(seed)
    v = dipping_layers(nz, nx, int(rng.integers(4, 8)), 1500.0,
                       float(rng.uniform(3500.0, 4500.0)), seed)
    k = nz / 301.0
    for _ in range(int(rng.integers(0, 4))):
        gaussian_lens(v, rng.uniform(0.25, 0.85) * nz, rng.uniform(0.1, 0.9) * nx,
                      rng.uniform(10.0, 25.0) * k, float(rng.choice([-300.0, 300.0])))
    if seed % 2 == 0:
        salt_polygon(v, rng.uniform(0.45, 0.7) * nz, rng.uniform(0.25, 0.75) * nx,
                     0.1 * nz, 0.08 * nx, 4500.0, seed + 11)
    np.clip(v, 1500.0, 4500.0, out=v)
    return v, smooth_slowness(v, sigma if sigma else 20.0 * k)


def config5_models(seed=5, nz=401, nx=1601, sigma=None):
    """Comparison model: layered 1500-4500 m/s, two salt-like bodies and five
    Gaussian lenses, all seeded; starting model = smoothed slowness (sigma 40
    cells at nz = 401, scaled with nz)."""
    rng = np.random.default_rng(seed)
    k = nz / 401.0
    v = flat_layers(nz, nx, list(np.linspace(1500.0, 4500.0, 7)))
    for j in range(5):
        gaussian_lens(v, rng.uniform(0.2, 0.8) * nz, rng.uniform(0.1, 0.9) * nx,
                      rng.uniform(15.0, 30.0) * k, 300.0 if j % 2 == 0 else -300.0)
    for j in range(2):
        salt_polygon(v, rng.uniform(0.45, 0.7) * nz, (0.3 + 0.4 * j) * nx,
                     0.09 * nz, 0.07 * nx, 4500.0, seed + 17 * (j + 1))
    np.clip(v, 1500.0, 4500.0, out=v)
    return v, smooth_slowness(v, sigma if sigma else 40.0 * k)


def mute_direct_wave(data, xs, zs, xr, zr, dt, velocity, delay, gstd=30.0, normalise=True):
    """Direct-wave mute of one shot (run/generate_input_data.py:22-46):
    data[nrec, nt] (t fastest) is divided by its max |.| and the samples
    before t_r = (distance/velocity + delay)/dt get the Gaussian taper
    exp(-0.5 ((t - t_r)/gstd)^2)."""
    d = np.array(data, np.float32, copy=True)
    if normalise:
        m = np.max(np.abs(d))
        if m > 0:
            d *= 1.0 / m
    nt = d.shape[1]
    t = np.arange(nt)
    dist = np.sqrt((xr - xs) ** 2 + (zr - zs) ** 2)
    tid = ((dist / velocity + delay) / dt).astype(int)
    for r in range(d.shape[0]):
        k = min(max(tid[r], 0), nt)
        d[r, :k] *= np.exp(-0.5 * ((t[:k] - tid[r]) / gstd) ** 2)
    return d


def write_parameters(path, s, model_file, sps, rps, folders, max_iteration=5, depth_mask=0.0,
                     mig_prefix="", inv_prefix="", extra=None):
    """parameters.txt in the reference's key = value format
    (run/parameters.txt keys; Appendix of the key list in the docs)."""
    lines = [
        f"x_samples = {s.nx}", f"z_samples = {s.nz}", f"model_spacing = {s.dh}",
        f"model_file = {model_file}", f"SPS = {sps}", f"RPS = {rps}",
        f"time_samples = {s.nt}", f"time_spacing = {s.dt}", f"max_frequency = {s.fmax}",
        f"abc_length = {(s.abc_nb - 1) * s.dh + 0.5 * s.dh}", f"abc_factor = {s.abc_factor}",
        f"mod_output_folder = {folders['data']}",
        f"max_iteration = {max_iteration}", f"depth_mask = {depth_mask}",
        f"inv_rbc_ratio = 200", f"inv_rbc_varVp = {s.varVp}",
        f"inv_rbc_varVp_rel = {s.varVp_rel}",
        f"inv_rbc_length = {(s.nb - 1) * s.dh + 0.5 * s.dh}",
        f"inv_input_folder = {folders['data']}", f"inv_input_prefix = {inv_prefix}",
        f"inv_output_folder = {folders['models']}", f"inv_residuo_folder = {folders['residuo']}",
        f"mig_rbc_ratio = 200", f"mig_rbc_varVp = {s.varVp}",
        f"mig_rbc_varVp_rel = {s.varVp_rel}",
        f"mig_rbc_length = {(s.nb - 1) * s.dh + 0.5 * s.dh}",
        f"mig_input_folder = {folders['data']}", f"mig_input_prefix = {mig_prefix}",
        f"mig_output_folder = {folders['seismic']}", f"rbc_seed = {s.seed}",
        "rbc_kind = uniform",           # the configurations' boundary (seeded +-40 % ramp)
    ]
    extra = dict(extra or {})
    # a key of `extra` replaces the line of the same key (the parser reads the first)
    lines = [l for l in lines if l.split(" = ")[0] not in extra]
    for k, v in extra.items():
        lines.append(f"{k} = {v}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_geometry(sps_path, rps_path, xs, zs, xr, zr):
    """SPS / RPS text files, "x,z" per line with %.2f (generate_setup.py:35-36)."""
    np.savetxt(sps_path, np.stack([xs, zs], 1), fmt="%.2f", delimiter=",")
    np.savetxt(rps_path, np.stack([xr, zr], 1), fmt="%.2f", delimiter=",")


def ensure_dirs(root):
    folders = {k: os.path.join(root, k) + "/" for k in ("data", "models", "residuo", "seismic")}
    for p in folders.values():
        os.makedirs(p, exist_ok=True)
    return folders


# ------------------------------------------------------------ command line
#
#   python -m fwi2d_amd.configs setup NAME DIR [--seed S] [--depth-mask M] [--max-iteration N]
#   python -m fwi2d_amd.configs init parameters.txt [--sigma 3] [--sources X0,X1,N,Z]
#                                                   [--receivers X0,X1,N,Z]
#   python -m fwi2d_amd.configs mute parameters.txt [--velocity 3600] [--delay 0.18] [--gstd 30]

def named_models(name, seed=None):
    """(survey, true model, smoothed starting model) of a named configuration."""
    if name == CONFIG1.name:
        v = config1_model(1 if seed is None else seed)
        return CONFIG1, v, smooth_slowness(v, 10.0)
    makers = {CONFIG2.name: (CONFIG2, config2_models, 2), CONFIG3.name: (CONFIG3, config3_models, 3),
              CONFIG5.name: (CONFIG5, config5_models, 5)}
    if name not in makers:
        raise SystemExit(f"unknown configuration '{name}' "
                         f"(one of {', '.join([CONFIG1.name] + list(makers))})")
    s, make, default = makers[name]
    true, init = make(default if seed is None else seed)
    return s, true, init


def seismogram_prefix(nt, nrec, dt):
    """File prefix of one shot's seismogram as the modeling executable writes it:
    seismogram_nt{nt}_nr{nrec}_{int(1e6f * dt)}us_shot_ (dt as a float32)."""
    return f"seismogram_nt{nt}_nr{nrec}_{int(np.float32(1e6) * np.float32(dt))}us_shot_"


def setup(name, root, seed=None, depth_mask=0.0, max_iteration=5):
    """Writes everything the three executables read for a named configuration:
    the true and starting models (models/<name>_{true,init}_{nz}x{nx}_{dh}m.bin;
    the starting model keeps the rows above depth_mask from the true one, as
    run/generate_setup.py:47-51 does), SPS/RPS and parameters_{modeling,
    migration,inversion}.txt.  Migration reads the muted input_RTM_ files that
    the `mute` command writes; inversion reads the raw seismograms.  Returns
    the three parameter file paths."""
    s, true, init = named_models(name, seed)
    folders = ensure_dirs(root)
    keep = int(depth_mask / s.dh)
    if keep > 0:
        init[:, :keep] = true[:, :keep]
    tag = f"{s.nz}x{s.nx}_{int(s.dh)}m.bin"
    true_file = os.path.join(folders["models"], f"{name}_true_{tag}")
    init_file = true_file.replace("_true_", "_init_")
    true.astype(np.float32).tofile(true_file)
    init.astype(np.float32).tofile(init_file)
    sps, rps = os.path.join(root, "SPS.txt"), os.path.join(root, "RPS.txt")
    write_geometry(sps, rps, *s.acquisition())
    prefix = seismogram_prefix(s.nt, s.nrec, s.dt)
    out = []
    for step, model in (("modeling", true_file), ("migration", init_file),
                        ("inversion", init_file)):
        p = os.path.join(root, f"parameters_{step}.txt")
        write_parameters(p, s, model, sps, rps, folders, max_iteration=max_iteration,
                         depth_mask=depth_mask, mig_prefix="input_RTM_" + prefix,
                         inv_prefix=prefix)
        out.append(p)
    return out


def _line(text):
    x0, x1, n, z = text.split(",")
    n = int(n)
    xs = np.linspace(float(x0), float(x1), n)
    return xs, np.full(n, float(z))


def initial_model(parameters, sigma=3.0, sources=None, receivers=None):
    """run/generate_setup.py:23-53 on the files a parameters.txt names: writes
    1 / gaussian(1 / model, sigma cells) with the rows above depth_mask copied
    from the model, to model_file with "true" replaced by "init"; when
    `sources` / `receivers` ("x0,x1,n,z") are given, first writes SPS / RPS
    with n evenly spaced points from x0 to x1 at depth z.  Returns the
    starting model's path."""
    from .viewer import read_parameter, stoi
    model_file = read_parameter(parameters, "model_file")
    if "true" not in os.path.basename(model_file):
        raise SystemExit(f"model_file '{model_file}' has no 'true' in its name: the "
                         "starting model would overwrite it")
    nx = stoi(read_parameter(parameters, "x_samples"))
    nz = stoi(read_parameter(parameters, "z_samples"))
    dh = float(np.float32(read_parameter(parameters, "model_spacing")))
    zmask = float(np.float32(read_parameter(parameters, "depth_mask", "0")))
    v = np.fromfile(model_file, np.float32, count=nx * nz)
    if v.size != nx * nz:
        raise SystemExit(f"{model_file}: {v.size} floats, expected {nx} x {nz}")
    v = v.reshape(nx, nz)
    lines = [(key, _line(spec)) for key, spec in (("SPS", sources), ("RPS", receivers)) if spec]
    for key, (x, z) in lines:       # every input checked before any file is written
        np.savetxt(read_parameter(parameters, key), np.stack([x, z], 1), fmt="%.2f",
                   delimiter=",")
    init = smooth_slowness(v, sigma, keep_rows=int(zmask / dh))
    out = os.path.join(os.path.dirname(model_file),
                       os.path.basename(model_file).replace("true", "init"))
    init.tofile(out)
    return out


def mute(parameters, velocity=3600.0, delay=0.18, gstd=30.0):
    """run/generate_input_data.py:22-46 on the modeling executable's output:
    each shot's seismogram in mod_output_folder is divided by its max |.|,
    the samples before (distance / velocity + delay) / dt get the Gaussian
    taper, and the result is written beside it as input_RTM_<file>.  Returns
    the written paths."""
    from .viewer import read_parameter, stoi
    nt = stoi(read_parameter(parameters, "time_samples"))
    dt = float(read_parameter(parameters, "time_spacing"))
    folder = read_parameter(parameters, "mod_output_folder", "")
    sps = np.loadtxt(read_parameter(parameters, "SPS"), np.float32, delimiter=",", ndmin=2)
    rps = np.loadtxt(read_parameter(parameters, "RPS"), np.float32, delimiter=",", ndmin=2)
    nr = len(rps)
    prefix = seismogram_prefix(nt, nr, dt)
    srcs = [f"{folder}{prefix}{k + 1}.bin" for k in range(len(sps))]
    for src in srcs:                # every input checked before any file is written
        if not os.path.exists(src):
            raise FileNotFoundError(f"{src} does not exist (run the modeling executable first)")
        if os.path.getsize(src) != 4 * nr * nt:
            raise SystemExit(f"{src}: {os.path.getsize(src) // 4} floats, expected {nr} x {nt}")
    out = []
    for k, src in enumerate(srcs):
        data = np.fromfile(src, np.float32)
        d = mute_direct_wave(data.reshape(nr, nt), sps[k, 0], sps[k, 1], rps[:, 0], rps[:, 1],
                             dt, float(velocity), float(delay), float(gstd))
        dst = f"{folder}input_RTM_{prefix}{k + 1}.bin"
        d.astype(np.float32).tofile(dst)
        out.append(dst)
    return out


def main(argv=None):
    import argparse
    p = argparse.ArgumentParser(prog="python -m fwi2d_amd.configs",
                                description="input files of the modeling, migration and "
                                            "inversion executables")
    sub = p.add_subparsers(dest="cmd", required=True)
    ps = sub.add_parser("setup", help="models, geometry and parameter files of a named "
                                      "configuration")
    ps.add_argument("name", help=", ".join(s.name for s in (CONFIG1, CONFIG2, CONFIG3, CONFIG5)))
    ps.add_argument("dir")
    ps.add_argument("--seed", type=int, default=None)
    ps.add_argument("--depth-mask", type=float, default=0.0)
    ps.add_argument("--max-iteration", type=int, default=5)
    pi = sub.add_parser("init", help="smoothed-slowness starting model of model_file")
    pi.add_argument("parameters")
    pi.add_argument("--sigma", type=float, default=3.0, help="Gaussian width in cells")
    pi.add_argument("--sources", default=None, help="x0,x1,n,z: rewrite SPS")
    pi.add_argument("--receivers", default=None, help="x0,x1,n,z: rewrite RPS")
    pm = sub.add_parser("mute", help="direct-wave-muted migration inputs (input_RTM_ files)")
    pm.add_argument("parameters")
    pm.add_argument("--velocity", type=float, default=3600.0)
    pm.add_argument("--delay", type=float, default=0.18)
    pm.add_argument("--gstd", type=float, default=30.0)
    a = p.parse_args(argv)
    if a.cmd == "setup":
        written = setup(a.name, a.dir, a.seed, a.depth_mask, a.max_iteration)
    elif a.cmd == "init":
        written = [initial_model(a.parameters, a.sigma, a.sources, a.receivers)]
    else:
        written = mute(a.parameters, a.velocity, a.delay, a.gstd)
    for w in written:
        print(w)
    return written


if __name__ == "__main__":
    main()
