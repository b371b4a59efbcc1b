"""Seeded synthetic cases shared by the tests (models, geometry, wavelets)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def layered(nz, nx, speeds, anomaly=None, seed=0):
    """[nx, nz] float32 model with len(speeds) flat layers and an optional
    Gaussian anomaly (sigma cells, amplitude m/s) at a seeded position."""
    v = np.zeros((nx, nz), np.float32)
    nl = len(speeds)
    for k, s in enumerate(speeds):
        v[:, k * nz // nl:] = s
    if anomaly is not None:
        sigma, amp = anomaly
        rng = np.random.default_rng(seed)
        zc = rng.uniform(0.3, 0.7) * nz
        xc = rng.uniform(0.3, 0.7) * nx
        z = np.arange(nz)[None, :]
        x = np.arange(nx)[:, None]
        v += (amp * np.exp(-0.5 * ((z - zc) ** 2 + (x - xc) ** 2) / sigma ** 2)).astype(np.float32)
    return v


def overthrust():
    """The reference's 81 x 201 overthrust model (25 m), [nx, nz]."""
    v = np.fromfile(os.path.join(GOLDEN, "overthrust_true_81x201_25m.bin"), np.float32)
    return v.reshape(201, 81)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
