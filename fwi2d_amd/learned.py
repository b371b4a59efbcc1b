"""Learned-preconditioner FWI: sample generation, training and the
comparison runs, over the engine, the network and the C++ model update.

The FWI iteration itself is C++ (include/fwi2d_update.h, used by the
inversion executable as well): the gradient pass and its post-processing
(fwi_gradient, fwi_get_gradient), then fwi_update_model -- the reference's
two-point parabolic line search and guarded step (inversion.cu:286-358) and,
for the trial runs, a second direction d tried at multiples of the plain
step and taken only when its forward-modeled misfit is below the
forward-modeled misfit of the plain update.  This module only orchestrates:

* ``make_samples`` runs plain FWI on seeded synthetic models and records,
  per iteration k, the input (raw gradient, illumination, model m_k) and the
  accepted updates that the plain run still makes after iteration k,
  m_end - m_k.  A single accepted update -step*g is collinear with the
  post-processed gradient g, so it would only teach the network the
  post-processing; the accumulated updates carry what the remaining
  iterations add.
* The network is a learned correction of the plain direction: the trial
  direction is d = -g + net(x) (fwi_trial_direction), and the training target
  is the normalised accumulated update minus -g.  A network that outputs zeros
  gives d = -g, the control (FWI_TRIAL_PLAIN).
* ``train`` fits the network (BF16 MFMA kernels, Adam) to those samples; every
  input channel and the accumulated update are normalised by their maximum
  magnitude per sample.  Both sample generation (a file per training model)
  and training (a file rewritten every epoch) can checkpoint and resume.
* ``run_fwi`` runs plain FWI or FWI with a trial direction over a multiscale
  band schedule and records the misfit trajectory.

Multiscale bands rebuild the wavelet and re-model the observed data per band
from the true model (a synthetic inverse-crime setup).
"""
import dataclasses
import json
import os
import time

import numpy as np

from . import abi as A
from . import configs as C


@dataclasses.dataclass(frozen=True)
class FwiSetup:
    nz: int
    nx: int
    dh: float
    nt: int
    dt: float
    nsrc: int
    nrec: int
    zsrc: float
    zrec: float
    rbc_nb: int
    abc_nb: int
    varVp_rel: float
    seed: int
    peaks: tuple            # Ricker peak of each band [Hz]; max_frequency = 3 x peak
    band_iters: tuple       # iterations per band
    zmask: float            # gradient depth mask [m]

    @property
    def iterations(self):
        return int(sum(self.band_iters))

    @property
    def abc_factor(self):
        return 0.35 / self.abc_nb        # edge factor exp(-0.35^2) for any pad width

    def band(self, it):
        acc = 0
        for b, n in enumerate(self.band_iters):
            acc += n
            if it < acc:
                return b
        return len(self.band_iters) - 1

    def acquisition(self):
        xmax = (self.nx - 1) * self.dh
        xs = np.linspace(0.02 * xmax, 0.98 * xmax, self.nsrc).astype(np.float32)
        xr = np.linspace(0.0, xmax, self.nrec).astype(np.float32)
        return (xs, np.full(self.nsrc, self.zsrc, np.float32), xr,
                np.full(self.nrec, self.zrec, np.float32))


# Reduced-size versions of the training (config 4) and comparison (config 5)
# workloads: same model families, bands and rules, smaller grids and shot
# counts so that generation, training and both comparison runs fit in minutes.
TRAIN_SMALL = FwiSetup(nz=128, nx=384, dh=10.0, nt=1875, dt=8e-4, nsrc=16, nrec=192,
                       zsrc=20.0, zrec=10.0, rbc_nb=60, abc_nb=60, varVp_rel=0.4, seed=4,
                       peaks=(4.0, 8.0, 12.0), band_iters=(3, 3, 2), zmask=60.0)
COMPARE_SMALL = dataclasses.replace(TRAIN_SMALL, nsrc=32, nrec=384, seed=5,
                                    band_iters=(7, 7, 6))
# Training at half the full size (200 x 800 at 10 m, 64 shots): the model
# families scale their structures with the grid, so these gradients have
# features closer in cells to the full-size comparison's.
TRAIN_MID = FwiSetup(nz=200, nx=800, dh=10.0, nt=3126, dt=8e-4, nsrc=64, nrec=400,
                     zsrc=20.0, zrec=10.0, rbc_nb=100, abc_nb=80, varVp_rel=0.4, seed=4,
                     peaks=(4.0, 8.0, 12.0), band_iters=(3, 3, 2), zmask=60.0)
# The training configuration at its stated size (256 x 768 at 10 m, 96
# shots, 8 plain iterations per model over the 4 / 8 / 12 Hz bands); the
# record length and pads follow the 200 x 800 setup (2.5 s at 0.8 ms).
TRAIN_FULL = FwiSetup(nz=256, nx=768, dh=10.0, nt=3126, dt=8e-4, nsrc=96, nrec=768,
                      zsrc=20.0, zrec=10.0, rbc_nb=100, abc_nb=80, varVp_rel=0.4, seed=4,
                      peaks=(4.0, 8.0, 12.0), band_iters=(3, 3, 2), zmask=60.0)
# The comparison configuration at full size: 401 x 1601 at 10 m, 200-cell
# random boundary, 256 shots, 1601 receivers, nt 5001 at 0.6 ms, bands
# 4 / 8 / 12 Hz (20 iterations: 7 / 7 / 6).
COMPARE_FULL = FwiSetup(nz=401, nx=1601, dh=10.0, nt=5001, dt=6e-4, nsrc=256, nrec=1601,
                        zsrc=20.0, zrec=10.0, rbc_nb=200, abc_nb=100, varVp_rel=0.4, seed=5,
                        peaks=(4.0, 8.0, 12.0), band_iters=(7, 7, 6), zmask=60.0)


def velocity(m):
    return (1.0 / np.sqrt(m)).astype(np.float32)


def slowness2(v):
    v = np.asarray(v, np.float64)
    return (1.0 / (v * v)).astype(np.float32)


_TUNED = {}        # launch shapes per setup (autotuned once per process)


class Fwi:
    """One survey on one engine: observed data per band, the gradient pass,
    the misfit of a model and the C++ model update."""

    def __init__(self, setup, v_true, v_init, autotune=True):
        s = self.setup = setup
        self.eng = A.Engine(s.nz, s.nx, s.dh, s.nt, s.dt, abc_nb=s.abc_nb,
                            abc_factor=s.abc_factor, rbc_nb=s.rbc_nb, rbc_varVp=0.0,
                            rbc_varVp_rel=s.varVp_rel, rbc_seed=s.seed, max_batch=s.nsrc)
        self.eng.set_geometry(*s.acquisition())
        self.eng.set_model(v_true)
        if autotune:
            # launch shapes depend on the grid and the batch only: the first
            # survey of a setup times them, later ones reuse them
            key = _setup_tag(s)
            if key not in _TUNED:
                w, tlag = A.ricker_wavelet(s.nt, s.dt, 3.0 * s.peaks[-1])
                self.eng.set_wavelet(w, tlag)
                self.eng.autotune(s.nsrc, reps=4)
                _TUNED[key] = [self.eng.get_tuning(k)[:3] for k in range(3)]
            for k, (st, vz, xw) in enumerate(_TUNED[key]):
                self.eng.set_tuning(vz, xw, kernel=k, steps=st)
        # observed data of every band the schedule uses, modeled from the true
        # model before any run starts (a band a run adds later is modeled
        # when it is first used)
        self.bands = [None] * len(s.peaks)
        for b, n in enumerate(s.band_iters):
            if n > 0:
                self._model_band(b)
        self.v_true = np.asarray(v_true, np.float32)
        self.m0 = slowness2(v_init)
        # line-search trial steps of the reference (inversion.cu:288-301)
        self.dm = float(1.0 / float(np.min(v_init)) ** 2 - 1.0 / float(np.max(v_init)) ** 2)
        self.alphas = (0.02 * self.dm, 0.05 * self.dm)
        vstable = 0.999 * 0.5546 * s.dh / s.dt     # fastest velocity the propagator accepts
        self.m_min = 1.0 / (vstable * vstable)
        self.current = None

    def close(self):
        self.eng.close()

    def _model_band(self, b):
        """Wavelet and observed data of band b (the engine holds the true
        model while this runs)."""
        w, tlag = A.ricker_wavelet(self.setup.nt, self.setup.dt, 3.0 * self.setup.peaks[b])
        self.eng.set_wavelet(w, tlag)
        self.bands[b] = (w, tlag, self.eng.modeling(A.ABC))
        self.current = None

    def use_band(self, b):
        if self.bands[b] is None:
            self.eng.set_model(self.v_true)
            self._model_band(b)
        if b != self.current:
            w, tlag, obs = self.bands[b]
            self.eng.set_wavelet(w, tlag)
            self.eng.set_observed(obs)
            self.current = b

    def misfit(self, m):
        """Misfit of model m, +inf when m cannot be propagated."""
        if not np.all((m > 0) & (m >= np.float32(self.m_min))):
            return float("inf")
        self.eng.set_model(velocity(m))
        return float(np.sqrt(self.eng.misfit()))

    def gradient(self, m, it):
        self.eng.set_model(velocity(m))
        self.eng.reset(A.RESET_GRADIENT)
        s2 = self.eng.gradient(iteration=it)
        g, raw, ill, _ = self.eng.get_gradient(self.setup.zmask)
        return float(np.sqrt(s2)), g, raw, ill

    def update(self, m, g, f0, trial=A.TRIAL_NONE, net=None, direction=None,
               multiples=(2.0, 4.0), raw=None, ill=None):
        """fwi_update_model: (new m, info) of the plain update, or of the
        trial direction when it passes the acceptance rule."""
        return self.eng.update_model(m, g, f0, self.alphas[0], self.alphas[1], self.m_min,
                                     zmask=self.setup.zmask, trial=trial, net=net,
                                     direction=direction, multiples=multiples, raw=raw,
                                     illum=ill)


def normalise_inputs(raw, ill, m):
    x = np.stack([raw, ill, m]).astype(np.float32)
    for c in range(3):
        s = float(np.max(np.abs(x[c])))
        if s > 0:
            x[c] /= s
    return x


def network(nx, nz, params=None, seed=0):
    """The network for [nx, nz] models: its image is the model grid edge-padded
    to multiples of 4 (its two 2x2 pools), the shape fwi_trial_direction
    expects."""
    net = A.Net(1, nx + (-nx) % 4, nz + (-nz) % 4, seed)
    if params is not None:
        net.set_params(params)
    return net


RUNS = {"plain": A.TRIAL_NONE, "plain_short": A.TRIAL_NONE, "control": A.TRIAL_PLAIN,
        "learned": A.TRIAL_NETWORK}


def run_fwi(fwi, trial=A.TRIAL_NONE, net=None, direction=None, multiples=(2.0, 4.0),
            band_iters=None, log=None):
    """Plain FWI (trial FWI_TRIAL_NONE) or FWI with a trial direction and the
    acceptance rule (FWI_TRIAL_PLAIN: the control d = -g; FWI_TRIAL_NETWORK:
    d = -g + net(x); FWI_TRIAL_GIVEN: d = direction(raw, ill, m, g), a
    callable).  band_iters overrides the setup's iterations per band.

    Returns (history, trajectory, final model).  history has one entry per
    iteration (the misfit f0 of its starting model, the update's record, wall
    time, model error) and a last entry with the final model's misfit.  The
    trajectory lists every forward-modeled misfit of a model the run adopts,
    with the wall time at which it was known: f0 after each gradient pass,
    the adopted model's misfit as soon as the update has modeled it (the
    trial runs always do; the plain run when its step is a line-search
    trial), and the final misfit.  time_to_misfit reads it."""
    s = fwi.setup if band_iters is None else dataclasses.replace(fwi.setup,
                                                                 band_iters=tuple(band_iters))
    its = s.iterations
    m = fwi.m0.copy()
    fwi.eng.reset()                       # illumination accumulates from here on
    hist, traj = [], []
    rms = lambda mm: float(np.sqrt(np.mean((velocity(mm) - fwi.v_true) ** 2)))
    t0 = time.perf_counter()
    for it in range(its):
        b = s.band(it)
        fwi.use_band(b)
        f0, g, raw, ill = fwi.gradient(m, it)
        traj.append({"time": time.perf_counter() - t0, "misfit": f0, "band": b})
        d = None
        if trial == A.TRIAL_GIVEN:
            d = direction(raw, ill, m, g)
        m, info = fwi.update(m, g, f0, trial=trial, net=net, direction=d, multiples=multiples,
                             raw=raw, ill=ill)
        t = time.perf_counter() - t0
        known = None
        if info["taken"]:
            known = info["trial_misfit"]
        elif trial != A.TRIAL_NONE:
            known = info["plain_misfit"]
        elif info["rule"] == 1:
            known = info["f"][1] if info["step"] == fwi.alphas[0] else info["f"][2]
        elif info["rule"] == 2:
            known = f0
        if known is not None and np.isfinite(known):
            traj.append({"time": t, "misfit": float(known), "band": b})
        hist.append(dict(info, iteration=it, band=b, misfit=f0, time=t,
                         taken="trial" if info["taken"] else "plain", model_rms_error=rms(m)))
        if log:
            log(hist[-1])
    b = s.band(its - 1)
    fwi.use_band(b)
    fend = fwi.misfit(m)
    t = time.perf_counter() - t0
    traj.append({"time": t, "misfit": fend, "band": b})
    hist.append({"iteration": its, "band": b, "misfit": fend, "time": t, "taken": "end",
                 "model_rms_error": rms(m)})
    return hist, traj, m


def _save_npz(path, **arrays):
    """np.savez to path through a temporary file, so an interrupted write
    leaves either the old file or the new one."""
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        np.savez(f, **arrays)
    os.replace(tmp, path)


def _setup_tag(setup):
    return json.dumps(dataclasses.asdict(setup), sort_keys=True)


def make_samples(setup, seeds, log=None, cache=None):
    """Plain-FWI samples: X [N, 3, nx, nz] (normalised raw gradient,
    illumination, model), Y [N, nx, nz] (the network's target: normalised
    m_end - m_k plus the plain direction g), G [N, nx, nz] (g) and the
    per-sample scale max|m_end - m_k| / dm.

    With `cache` (a folder), each seed's samples are written to
    cache/samples_<seed>.npz as soon as its run ends, and a seed whose file
    exists is read instead of run again, so an interrupted generation resumes
    at the first missing seed.  A file made for another setup is refused."""
    xs, ys, gs, scales = [], [], [], []
    shape = (setup.nx, setup.nz)
    if cache:
        os.makedirs(cache, exist_ok=True)
    for seed in seeds:
        path = os.path.join(cache, f"samples_{seed}.npz") if cache else None
        cached = bool(path) and os.path.exists(path)
        if cached:
            with np.load(path) as z:
                if str(z["setup"]) != _setup_tag(setup):
                    raise ValueError(f"{path} holds samples of another setup")
                sx, sy, sg, ss = z["x"], z["y"], z["g"], z["scale"]
        else:
            sx, sy, sg, ss = _seed_samples(setup, seed)
            if path:
                _save_npz(path, x=sx, y=sy, g=sg, scale=ss, setup=np.array(_setup_tag(setup)))
        xs.extend(sx)
        ys.extend(sy)
        gs.extend(sg)
        scales.extend(ss)
        if log:
            log({"seed": seed, "samples": len(xs), "cached": cached})
    if not xs:
        return (np.zeros((0, 3) + shape, np.float32), np.zeros((0,) + shape, np.float32),
                np.zeros((0,) + shape, np.float32), np.zeros(0))
    return np.stack(xs), np.stack(ys), np.stack(gs), np.array(scales)


def _seed_samples(setup, seed):
    """Samples of one training model (see make_samples), as arrays."""
    v, v0 = C.config4_models(seed, setup.nz, setup.nx)
    fwi = Fwi(setup, v, v0)
    m = fwi.m0.copy()
    fwi.eng.reset()
    inputs = []
    for it in range(setup.iterations):
        fwi.use_band(setup.band(it))
        f0, g, raw, ill = fwi.gradient(m, it)
        inputs.append((raw, ill, m.copy(), g))
        m, _ = fwi.update(m, g, f0)
    sx, sy, sg, ss = [], [], [], []
    for raw, ill, mk, gk in inputs:
        t = (m - mk).astype(np.float32)
        s = float(np.max(np.abs(t)))
        if s <= 0:
            continue
        sx.append(normalise_inputs(raw, ill, mk))
        sy.append(t / s + gk)
        sg.append(gk)
        ss.append(s / fwi.dm)
    fwi.close()
    shape = (setup.nx, setup.nz)
    if not sx:
        return (np.zeros((0, 3) + shape, np.float32), np.zeros((0,) + shape, np.float32),
                np.zeros((0,) + shape, np.float32), np.zeros(0))
    return np.stack(sx), np.stack(sy), np.stack(sg), np.array(ss)


def train(x, y, epochs=40, batch=16, lr=1e-3, seed=0, holdout=0.1, flip=True, log=None,
          checkpoint=None, init=None, state_out=None):
    """Adam on the GPU; returns (params, history of (epoch, train loss,
    held-out loss), held-out indices).  With flip, each training sample is
    mirrored along x with probability 1/2 (the wave equation is symmetric
    under x -> -x, so a mirrored sample is a valid sample).

    With `checkpoint` (a file path), the state after every epoch (parameters,
    Adam moments and step count, the shuffled order, the random generator and
    the loss history) is written there, and an existing checkpoint is resumed
    from the epoch after it: a run stopped and resumed gives the same
    parameters, bit for bit, as one that ran straight through.  A checkpoint
    of another training problem (sample count, image size, batch, seed,
    hold-out, mirroring, learning rate) is refused.

    With `init` (a dict: params, and optionally the Adam moments m, v and the
    step count steps of an earlier training), training starts from that state
    instead of a fresh network: a training set too large for one session is
    then learned in phases, one set of models per phase.  With `state_out`
    (a file path), the final parameters and Adam state are written there in
    the form `init` takes (load_state)."""
    n, _, h, w = x.shape
    rng = np.random.default_rng(seed)
    order = rng.permutation(n)
    nval = max(1, int(round(holdout * n)))
    val, tr = order[:nval], order[nval:]
    net = A.Net(batch, h, w, seed)
    if init is not None:
        net.set_params(init["params"])
        if "m" in init:
            net.set_optimizer(init["m"], init["v"], int(init["steps"]))
    hist = []
    first = 0
    problem = {"n": int(n), "h": int(h), "w": int(w), "batch": int(batch), "seed": int(seed),
               "holdout": float(holdout), "flip": bool(flip), "lr": float(lr)}
    if checkpoint and os.path.exists(checkpoint):
        with np.load(checkpoint) as z:
            meta = json.loads(str(z["meta"]))
            if meta["problem"] != problem:
                net.close()
                raise ValueError(f"{checkpoint} is a checkpoint of another training problem "
                                 f"({meta['problem']})")
            net.set_params(z["params"])
            net.set_optimizer(z["m"], z["v"], meta["steps"])
            tr = z["order"].copy()
        rng.bit_generator.state = meta["rng"]
        hist = [tuple(e) for e in meta["hist"]]
        first = meta["epoch"] + 1
        if log:
            log({"resumed_from": checkpoint, "epoch": meta["epoch"]})
    for ep in range(first, epochs):
        rng.shuffle(tr)
        losses = []
        for b0 in range(0, len(tr) - batch + 1, batch):
            idx = tr[b0:b0 + batch]
            xb, yb = x[idx], y[idx]
            if flip:
                f = rng.random(len(idx)) < 0.5
                xb, yb = xb.copy(), yb.copy()
                xb[f] = xb[f][:, :, ::-1, :]
                yb[f] = yb[f][:, ::-1, :]
            losses.append(net.train_step(xb, yb, lr))
        vl = []
        for b0 in range(0, nval, batch):
            idx = val[b0:b0 + batch]
            vl.append(net.gradient(x[idx], y[idx])[0] * len(idx))
        hist.append((ep, float(np.mean(losses)), float(np.sum(vl) / nval)))
        if log:
            log({"epoch": ep, "train_loss": hist[-1][1], "holdout_loss": hist[-1][2]})
        if checkpoint:
            m, v, steps = net.get_optimizer()
            meta = {"problem": problem, "epoch": ep, "steps": steps, "hist": hist,
                    "rng": rng.bit_generator.state}
            _save_npz(checkpoint, params=net.get_params(), m=m, v=v, order=tr,
                      meta=np.array(json.dumps(meta)))
    p = net.get_params()
    if state_out:
        m, v, steps = net.get_optimizer()
        _save_npz(state_out, params=p, m=m, v=v, steps=np.array(steps))
    net.close()
    return p, hist, val


def load_state(path):
    """Training state for train(init=...): a state file written by train
    (an npz archive: parameters and Adam state, whatever its name) or a bare
    parameter file (float32; Adam starts afresh)."""
    with open(path, "rb") as f:
        is_zip = f.read(4) == b"PK\x03\x04"
    if is_zip:
        with np.load(path) as z:
            return {k: z[k] for k in ("params", "m", "v", "steps")}
    return {"params": np.fromfile(path, np.float32)}


def holdout_cosines(params, x, y, g, idx):
    """Mean cosine between the accumulated update (y - g) and the learned
    direction net(x) - g, and between it and the plain direction -g, over
    samples idx."""
    n, _, h, w = x.shape
    net = A.Net(1, h, w, 0)
    net.set_params(params)
    cl, cp = [], []
    for i in idx:
        t = y[i] - g[i]
        cl.append(cosine(net.forward(x[i][None])[0] - g[i], t))
        cp.append(cosine(-g[i], t))
    net.close()
    return float(np.mean(cl)), float(np.mean(cp))


def cosine(a, b):
    a, b = a.ravel().astype(np.float64), b.ravel().astype(np.float64)
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300))


def time_to_misfit(traj, target, band):
    """Wall time at which a run's misfit trajectory (run_fwi) first reaches
    `target` in `band` (misfits of different bands measure different data
    and are not compared), or None."""
    for e in traj:
        if e["band"] == band and e["misfit"] <= target:
            return e["time"]
    return None


def summarise(runs, trajs):
    """Per run: final misfit (last band), model error, wall time, and the time
    at which its trajectory first reaches plain FWI's final misfit, also as a
    fraction of plain FWI's wall time."""
    ref = runs.get("plain")
    target = ref[-1]["misfit"] if ref else None
    band = ref[-1]["band"] if ref else None
    t_plain = ref[-1]["time"] if ref else None
    out = {}
    for name, h in runs.items():
        t = time_to_misfit(trajs[name], target, band) if ref else None
        out[name] = {"iterations": len(h) - 1, "final_misfit": h[-1]["misfit"],
                     "final_model_rms_error": h[-1]["model_rms_error"],
                     "time_s": h[-1]["time"], "time_to_plain_final_misfit_s": t,
                     "time_ratio": (t / t_plain) if (t is not None and t_plain) else None,
                     "trial_steps_taken": sum(e.get("taken") == "trial" for e in h)}
    return {"plain_final_misfit": target, "plain_time_s": t_plain, "runs": out}


def main(argv=None):
    import argparse
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--out", default="learned_run")
    p.add_argument("--train-models", type=int, default=48)
    p.add_argument("--train-size", choices=("small", "mid", "full"), default="small",
                   help="training survey: 128 x 384 (16 shots), 200 x 800 (64 shots) or the "
                        "stated 256 x 768 (96 shots)")
    p.add_argument("--train-first-seed", type=int, default=1)
    p.add_argument("--epochs", type=int, default=40)
    p.add_argument("--lr", type=float, default=1e-3)
    p.add_argument("--compare-seed", type=int, default=10005)
    p.add_argument("--params", default="", help="trained parameters (skips generation and training)")
    p.add_argument("--size", choices=("small", "full"), default="small",
                   help="comparison at the reduced size or at the full config-5 size")
    p.add_argument("--runs", default="plain,plain_short,control,learned",
                   help="comma-separated runs of the comparison (plain, plain_short, control, "
                        "learned); empty: train only")
    p.add_argument("--bands", default="", help="iterations per band of the plain run "
                                               "(default: the setup's, 7,7,6)")
    p.add_argument("--short-bands", default="",
                   help="iterations per band of the plain_short, control and learned runs "
                        "(default: --bands)")
    p.add_argument("--checkpoint", default="",
                   help="folder for resumable sample generation (one file per training "
                        "model) and training (one file, rewritten every epoch)")
    p.add_argument("--continue-from", default="",
                   help="train from this state (a train_state.npz of an earlier phase, or a "
                        "parameter file) instead of a fresh network")
    p.add_argument("--merge", nargs="+", default=None, metavar="REPORT",
                   help="summarise the runs of several report.json files of the same "
                        "comparison (runs made in separate sessions) and exit")
    a = p.parse_args(argv)
    log = lambda d: print(json.dumps(d), flush=True)
    if a.merge:
        runs, trajs = {}, {}
        for path in a.merge:
            with open(path) as f:
                r = json.load(f)
            runs.update(r["runs"])
            trajs.update(r["trajectories"])
        summary = summarise(runs, trajs)
        log(summary)
        return summary
    os.makedirs(a.out, exist_ok=True)

    if a.params:
        params = np.fromfile(a.params, np.float32)
        x, th, t_gen, t_train = np.zeros((0,)), [(0, None, None)], None, None
        cos = (None, None)
    else:
        t0 = time.perf_counter()
        ts = {"small": TRAIN_SMALL, "mid": TRAIN_MID, "full": TRAIN_FULL}[a.train_size]
        ck = a.checkpoint
        seeds = range(a.train_first_seed, a.train_first_seed + a.train_models)
        x, y, g, scales = make_samples(ts, seeds, log,
                                       cache=os.path.join(ck, f"samples_{a.train_size}")
                                       if ck else None)
        t_gen = time.perf_counter() - t0
        t0 = time.perf_counter()
        params, th, val = train(x, y, epochs=a.epochs, lr=a.lr, log=log,
                                checkpoint=os.path.join(ck, f"train_{a.train_size}.npz")
                                if ck else None,
                                init=load_state(a.continue_from) if a.continue_from else None,
                                state_out=os.path.join(a.out, "train_state.npz"))
        t_train = time.perf_counter() - t0
        cos = holdout_cosines(params, x, y, g, val)
        log({"holdout_cosine_learned": cos[0], "holdout_cosine_plain": cos[1]})
    params.tofile(os.path.join(a.out, "preconditioner_params.bin"))
    train_report = {"samples": int(len(x)), "generation_s": t_gen, "training_s": t_train,
                    "continued_from": a.continue_from or None,
                    "train_seeds": None if a.params else
                    [a.train_first_seed, a.train_first_seed + a.train_models - 1],
                    "holdout_cosine": {"learned": cos[0], "plain_direction": cos[1]},
                    "train_loss_first_last": [th[0][1], th[-1][1]],
                    "holdout_loss_first_last": [th[0][2], th[-1][2]]}
    if not a.runs:                    # training only
        with open(os.path.join(a.out, "report.json"), "w") as f:
            json.dump(train_report, f, indent=1)
        log(train_report)
        return

    s = COMPARE_SMALL if a.size == "small" else COMPARE_FULL
    long_bands = tuple(int(k) for k in a.bands.split(",")) if a.bands else s.band_iters
    short_bands = tuple(int(k) for k in a.short_bands.split(",")) if a.short_bands else long_bands
    v, v0 = C.config5_models(a.compare_seed, s.nz, s.nx)
    t0 = time.perf_counter()
    fwi = Fwi(s, v, v0)
    log({"setup_s": time.perf_counter() - t0, "grid": [s.nz, s.nx], "shots": s.nsrc,
         "bands": list(long_bands), "short_bands": list(short_bands)})
    net = network(s.nx, s.nz, params)
    runs, trajs = {}, {}
    for name in a.runs.split(","):
        if name not in RUNS:
            raise SystemExit(f"unknown run '{name}' (one of {', '.join(RUNS)})")
        bands = long_bands if name == "plain" else short_bands
        h, tr, _ = run_fwi(fwi, RUNS[name], net=net, band_iters=bands,
                           log=lambda e, n=name: log(dict(e, run=n)))
        runs[name], trajs[name] = h, tr
    net.close()
    fwi.close()
    summary = summarise(runs, trajs)
    report = dict(train_report, compare_seed=a.compare_seed, grid=[s.nz, s.nx], shots=s.nsrc,
                  bands=list(long_bands), short_bands=list(short_bands), summary=summary,
                  runs=runs, trajectories=trajs)
    with open(os.path.join(a.out, "report.json"), "w") as f:
        json.dump(report, f, indent=1)
    log({k: v for k, v in report.items() if k not in ("runs", "trajectories")})


if __name__ == "__main__":
    main()
