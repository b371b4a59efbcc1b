"""Writes the learned-preconditioner fixtures: networks trained by this
project's own kernels on plain-FWI samples.

    python tests/golden/make_preconditioner.py small   (GPU, ~4 minutes)
    python tests/golden/make_preconditioner.py mid     (GPU, ~8 minutes)
    python tests/golden/make_preconditioner.py mid96   (GPU, ~15 minutes)
    python tests/golden/make_preconditioner.py config4 (GPU, ~16 minutes)
    python tests/golden/make_preconditioner.py config4_256 (GPU, ~60 minutes)

small -> preconditioner_48models.bin: 48 seeded training models of the
config-4 family at 128 x 384 (10 m, 16 shots), 8 plain FWI iterations each
over bands 4/8/12 Hz (384 samples), 40 epochs of Adam at lr 1e-3, batch 16,
no mirroring (fwi2d_amd.learned.TRAIN_SMALL).  The small-size comparisons and
the GPU tests use it.

mid -> preconditioner_mid48.bin: the same 48 seeds at 200 x 800 (10 m,
64 shots; fwi2d_amd.learned.TRAIN_MID), 384 samples, 40 epochs, mirrored
samples (what `python -m fwi2d_amd.learned --train-size mid --runs ""`
writes).  The full-size (401 x 1601) comparison uses it: the model families
scale their structures with the grid, so mid-size gradients are closer in
cells to the full-size ones.

mid96 -> preconditioner_mid96.bin: the same at 200 x 800 with 96 training
models (seeds 1-96, 768 samples; `python -m fwi2d_amd.learned --train-size
mid --train-models 96 --runs ""`).

config4 -> preconditioner_config4.bin: the training configuration at its
stated grid and acquisition (256 x 768 at 10 m, 96 shots, 768 receivers,
8 plain FWI iterations per model over 4/8/12 Hz; fwi2d_amd.learned.TRAIN_FULL)
on 64 of its 256 models (seeds 1-64: what one 20-minute GPU session
generates), 40 epochs of Adam at lr 1e-3, batch 16, mirrored samples
(`python -m fwi2d_amd.learned --train-size full --train-models 64 --runs ""`).

config4_256 -> preconditioner_config4_256.bin: all 256 models of the
training configuration, learned in four phases of 64 models (seeds 1-64,
65-128, 129-192, 193-256), one GPU session each: each phase generates its
512 samples and runs 40 epochs of Adam at batch 16 from the previous phase's
network (`python -m fwi2d_amd.learned --train-size full --train-models 64
--train-first-seed S --continue-from STATE --runs ""`).  Phase 1 is the
config4 fixture; phase 2 started from its parameters with fresh Adam moments
(that fixture kept none), phases 3 and 4 from the full state of the phase
before.  5120 Adam steps in all, as 40 epochs over 2048 samples would take.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fwi2d_amd import learned as L  # noqa: E402


def phased(name):
    """config4_256: four phases of 64 models, as described above."""
    import tempfile
    here = os.path.dirname(os.path.abspath(__file__))
    state = None
    with tempfile.TemporaryDirectory() as tmp:
        for phase, first in enumerate((1, 65, 129, 193)):
            x, y, g, _ = L.make_samples(L.TRAIN_FULL, range(first, first + 64), log=print)
            path = os.path.join(tmp, f"phase{phase + 1}.npz")
            params, hist, val = L.train(x, y, epochs=40, lr=1e-3, flip=True, log=print,
                                        init=state, state_out=path)
            print(f"phase {phase + 1}: held-out cosine (learned, plain):",
                  L.holdout_cosines(params, x, y, g, val))
            # phase 2 starts from phase 1's parameters with fresh moments, the
            # later phases from the full state of the phase before
            state = {"params": params} if phase == 0 else L.load_state(path)
    params.tofile(os.path.join(here, name))


def main(size):
    if size == "config4_256":
        return phased("preconditioner_config4_256.bin")
    setup, flip, name, models = {
        "small": (L.TRAIN_SMALL, False, "preconditioner_48models.bin", 48),
        "mid": (L.TRAIN_MID, True, "preconditioner_mid48.bin", 48),
        "mid96": (L.TRAIN_MID, True, "preconditioner_mid96.bin", 96),
        "config4": (L.TRAIN_FULL, True, "preconditioner_config4.bin", 64)}[size]
    x, y, g, _ = L.make_samples(setup, range(1, models + 1), log=print)
    params, hist, val = L.train(x, y, epochs=40, lr=1e-3, flip=flip, log=print)
    print("held-out cosine (learned, plain):", L.holdout_cosines(params, x, y, g, val))
    params.tofile(os.path.join(os.path.dirname(os.path.abspath(__file__)), name))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "small")
