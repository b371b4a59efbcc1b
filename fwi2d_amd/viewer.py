"""Figures of a run's inputs and outputs, read from the files the three
executables read and write (the role of the reference's
run/generate_results.py:34-221): the model with the acquisition, one shot's
seismogram, the RTM section, the FWI convergence curve (residual as a
percentage of its largest value) and the final FWI model.

    python -m fwi2d_amd.viewer parameters.txt [--out DIR] [--shot K] [--true FILE]

Keys read from parameters.txt (the executables' key = value format):
x_samples, z_samples, model_spacing, time_samples, time_spacing,
max_frequency, model_file, SPS, RPS and, when present, mod_output_folder,
mig_output_folder, inv_output_folder, inv_residuo_folder, max_iteration.
A figure whose input file does not exist is skipped with a message.
"""
import argparse
import os
import re
import sys

import numpy as np


def read_parameter(path, key, default=None):
    """Value of `key` in a parameters.txt, with the executables' rules
    (fwi2d_amd/host/admin.cpp find_parameter): the key is the whole first
    token of a line that does not start with '#' or a blank, the value runs
    from '=' to an optional '#', blanks are removed, and a value that starts
    with '"' or '[' loses its quotes and brackets; a missing key or an empty
    value raises KeyError, or returns `default` when one is given."""
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line or line[0] in "# \t" or not line.startswith(key):
                continue
            rest = line[len(key):]
            if rest and rest[0] not in " \t=":
                continue
            eq = rest.find("=")
            if eq < 0:
                continue
            v = rest[eq + 1:].split("#", 1)[0]
            v = "".join(c for c in v if c not in " \t\r")
            if v[:1] in ('"', "["):
                v = "".join(c for c in v if c not in '"[]')
            if v:
                return v
            break                     # an empty value counts as missing
    if default is None:
        raise KeyError(f"parameter '{key}' is missing in {path}")
    return default


def stoi(text):
    """Leading integer of text, as std::stoi reads it."""
    m = re.match(r"\s*([+-]?\d+)", text)
    if not m:
        raise ValueError(f"no integer in '{text}'")
    return int(m.group(1))


def read_field(path, nz, nx):
    """[nx][nz] float32 file (z fastest) as an [nz, nx] image."""
    return np.fromfile(path, np.float32).reshape(nx, nz).T


def suffix(fmax, nz, nx, dh):
    return f"{int(fmax)}Hz_{nz}x{nx}_{int(dh)}m.bin"


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("parameters")
    p.add_argument("--out", default=".", help="folder for the PNG files")
    p.add_argument("--shot", type=int, default=0, help="shot (1-based); default the middle one")
    p.add_argument("--true", default="", help="true model file (default: model_file)")
    a = p.parse_args(argv)

    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    par = a.parameters
    nx = stoi(read_parameter(par, "x_samples"))
    nz = stoi(read_parameter(par, "z_samples"))
    dh = np.float32(read_parameter(par, "model_spacing"))
    nt = stoi(read_parameter(par, "time_samples"))
    dt = np.float32(read_parameter(par, "time_spacing"))
    fmax = np.float32(read_parameter(par, "max_frequency"))
    sps = np.loadtxt(read_parameter(par, "SPS"), np.float32, delimiter=",", ndmin=2)
    rps = np.loadtxt(read_parameter(par, "RPS"), np.float32, delimiter=",", ndmin=2)
    ns, nr = len(sps), len(rps)
    os.makedirs(a.out, exist_ok=True)
    ext = [0, (nx - 1) * dh, (nz - 1) * dh, 0]
    written = []

    def save(fig, name):
        fig.tight_layout()
        path = os.path.join(a.out, name)
        fig.savefig(path, dpi=150)
        plt.close(fig)
        written.append(path)

    def model_figure(path, name, vrange):
        v = read_field(path, nz, nx)
        fig, ax = plt.subplots(figsize=(10, 3))
        im = ax.imshow(v, cmap="jet", aspect="auto", extent=ext, vmin=vrange[0], vmax=vrange[1])
        ax.plot(rps[:, 0], rps[:, 1], "o", color="gray", markersize=2)
        ax.plot(sps[:, 0], sps[:, 1], "*", color="black")
        ax.set_xlabel("Distance [m]")
        ax.set_ylabel("Depth [m]")
        fig.colorbar(im, ax=ax, label="Velocity [m/s]")
        save(fig, name)

    def missing(path):
        if os.path.exists(path):
            return False
        print(f"skipped: {path} does not exist", file=sys.stderr)
        return True

    true_file = a.true or read_parameter(par, "model_file")
    vrange = None
    if not missing(true_file):
        v = read_field(true_file, nz, nx)
        vrange = (float(v.min()), float(v.max()))
        model_figure(true_file, "model.png", vrange)

    shot = a.shot if a.shot > 0 else ns // 2 + 1
    data_folder = read_parameter(par, "mod_output_folder", "")
    seis = (f"{data_folder}seismogram_nt{nt}_nr{nr}_{int(np.float32(1e6) * dt)}us_"
            f"shot_{shot}.bin")
    if data_folder and not missing(seis):
        d = np.fromfile(seis, np.float32).reshape(nr, nt).T
        s = float(np.std(d)) or 1.0
        fig, ax = plt.subplots(figsize=(6, 8))
        ax.imshow(d, aspect="auto", cmap="Greys", vmin=-s, vmax=s,
                  extent=[float(rps[0, 0] - sps[shot - 1, 0]), float(rps[-1, 0] - sps[shot - 1, 0]),
                          (nt - 1) * float(dt), 0])
        ax.set_xlabel("Offset [m]")
        ax.set_ylabel("Time [s]")
        ax.set_title(f"shot {shot}")
        save(fig, f"seismogram_shot_{shot}.png")

    mig_folder = read_parameter(par, "mig_output_folder", "")
    image = f"{mig_folder}RTM_section_{suffix(fmax, nz, nx, dh)}"
    if mig_folder and not missing(image):
        img = read_field(image, nz, nx)
        s = 2.0 * float(np.std(img)) or 1.0
        fig, ax = plt.subplots(figsize=(10, 3))
        ax.imshow(img, cmap="Greys", aspect="auto", extent=ext, vmin=-s, vmax=s)
        ax.set_xlabel("Distance [m]")
        ax.set_ylabel("Depth [m]")
        save(fig, "image_RTM.png")

    res_folder = read_parameter(par, "inv_residuo_folder", "")
    iters = read_parameter(par, "max_iteration", "")
    conv = f"{res_folder}convergence_{stoi(iters) if iters else ''}_iterations.txt"
    if res_folder and iters and not missing(conv):
        r = np.atleast_1d(np.loadtxt(conv, dtype=np.float64))
        r = 100.0 * r / np.max(r)
        fig, ax = plt.subplots(figsize=(10, 4))
        ax.plot(r, "ok--")
        ax.set_xlabel("Iteration")
        ax.set_ylabel("Residual L2 norm [% of max]")
        save(fig, "residuo_FWI.png")

    inv_folder = read_parameter(par, "inv_output_folder", "")
    final = f"{inv_folder}final_model_FWI_{suffix(fmax, nz, nx, dh)}"
    if inv_folder and not missing(final):
        model_figure(final, "model_FWI.png", vrange or (None, None))
    for w in written:
        print(w)
    return written


if __name__ == "__main__":
    main()
