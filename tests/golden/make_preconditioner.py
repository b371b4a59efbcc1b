"""Writes tests/golden/preconditioner_48models.bin: the learned preconditioner
trained by this project's own kernels on plain-FWI samples.

    python tests/golden/make_preconditioner.py      (needs the GPU, ~4 minutes)

48 seeded training models of the config-4 family at 128 x 384 (10 m),
8 plain FWI iterations each over bands 4/8/12 Hz (384 samples), 40 epochs of
Adam at lr 1e-3, batch 16, no mirroring (fwi2d_amd.learned.TRAIN_SMALL).
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fwi2d_amd import learned as L  # noqa: E402


def main():
    x, y, g, _ = L.make_samples(L.TRAIN_SMALL, range(1, 49), log=print)
    params, hist, val = L.train(x, y, epochs=40, lr=1e-3, flip=False, log=print)
    print("held-out cosine (learned, plain):", L.holdout_cosines(params, x, y, g, val))
    params.tofile(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                               "preconditioner_48models.bin"))


if __name__ == "__main__":
    main()
