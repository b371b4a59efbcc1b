"""CPU model of the launch geometry and global addressing of the three-step
step kernels (`fwd3_kernel`, `rev3_kernel` in fwi2d_amd/csrc/propagate.hip),
with the engine's buffer layout (`setup_grid`, `kGuard` in engine.hip);
odd x-chunks march right to left, their columns mirrored within the chunk.

Checks, for every tile and lane of a launch:
* the owned rows of the z segments and the owned columns of the x chunks
  cover the updated cells [4, nzz-4) x [4, nxx-4) exactly once;
* every load and store (register-queue columns, clamped to the grid; the
  halo rows, one float per lane (lanes 0-3 above the wave, 60-63 below it)
  loaded beside each queue column, whose column is clamped to [4, nxx-5];
  image and illumination of owned columns; stores of owned cells) falls
  inside its allocation [0, slots * fstride + 1024 guard floats), for the
  three- and four-step kernels.

The halo clamp is the case this pins: the first z segment starts at row -4,
so its halo rows -8 .. -5 lie in the previous column, which exists only for
columns >= 1 (clamped to >= 4 here)."""
import itertools

import numpy as np
import pytest

GUARD = 1024


def layout(nzz, nxx):
    ldz = (nzz + 3) & ~3
    fs = (ldz * nxx + 63) & ~63
    return ldz, fs


def zsegments(nzz, vz, steps):
    rows = 64 * vz - 8 * (steps - 1)
    return max(1, (nzz - 8 + rows - 1) // rows)


def tiles(nzz, nxx, vz, xw, steps=3):
    """(zs, zb[lanes], own[lanes], x0, x1, back) of each tile of a slot in a
    launch of `steps` steps; back: the chunk marches right to left (odd
    chunks)."""
    lanes = np.arange(64)
    ldz, _ = layout(nzz, nxx)
    h = 4 * (steps - 1)
    for zseg in range(zsegments(nzz, vz, steps)):
        zs = zseg * (64 * vz - 2 * h) + 4 - h          # owned from row 4
        zb = zs + lanes * vz
        own = (zb < ldz) & (zb >= zs + h) & (zb < zs + 64 * vz - h)
        for xc in range((nxx - 8 + xw - 1) // xw):
            x0 = 4 + xc * xw
            yield zs, zb, own, x0, min(x0 + xw, nxx - 4), bool(xc & 1)


def accesses(nzz, nxx, vz, xw, reverse, halo_lo=4, steps=3):
    """(buffer, lowest float index, highest float index) over one slot of the
    three- or (steps=4) four-step forward or reverse kernel.  An address is col * ldz + row, increasing in both, so each
    access kind is bounded by its extreme columns and rows."""
    ldz, _ = layout(nzz, nxx)
    lanes = np.arange(64)
    out = []
    cl = lambda c: min(max(c, 0), nxx - 1)
    ch = lambda c: min(max(c, halo_lo), nxx - 5)
    h = 4 * (steps - 1)        # the march covers the chunk +- h columns
    for zs, zb, own, x0, x1, back in tiles(nzz, nxx, vz, xw, steps):
        zl = np.maximum(zb, 0)
        hz1 = np.where(lanes < 32, zs - 4, zs + 64 * vz) + (lanes & 3)
        span = lambda c0, c1, rows, n: (c0 * ldz + rows.min(), c1 * ldz + rows.max() + n - 1)
        # the march's column range [a, b] as grid columns (mirrored for odd chunks)
        pc = lambda c: (x0 + x1 - 1 - c) if back else c
        cols = lambda a, b: (min(pc(a), pc(b)), max(pc(a), pc(b)))
        # queue columns: window [x0 - h, x1 + h) +- 4, the last loaded at j + 5
        c0, c1 = cols(x0 - h - 4, x1 + h + 4)
        out.append(("field",) + span(cl(c0), cl(c1), zl, vz))              # u_t / S_cur, R_cur
        c0, c1 = cols(x0 - h, x1 + h)
        out.append(("field",) + span(cl(c0), cl(c1), zl, vz))              # u_old
        out.append(("c",) + span(cl(c0), cl(c1), zl, vz))                 # coefficients
        c0, c1 = cols(x0 - h, x1 + h + 4)                                  # halo of each queue column
        out.append(("field",) + span(ch(c0), ch(c1), hz1, 1))              # one float per lane
        if reverse:
            out.append(("field",) + span(x0, x1 - 1, zl, vz))              # image, illum
        if own.any():
            out.append(("field",) + span(x0, x1 - 1, zb[own], vz))         # stores
    return out


SHAPES = [(601, 1301), (163, 283), (183, 303), (120, 180), (9, 9), (13, 40), (801, 2001)]


@pytest.mark.parametrize("nzz,nxx", SHAPES)
@pytest.mark.parametrize("vz", [1, 2, 4])
@pytest.mark.parametrize("steps", [3, 4])
def test_owned_cells_cover_the_grid_once(nzz, nxx, vz, steps):
    for xw in (3, 8, 81, 144, 1000):
        count = np.zeros((nxx, nzz), np.int32)
        for zs, zb, own, x0, x1, _ in tiles(nzz, nxx, vz, xw, steps):
            rows = (zb[own][:, None] + np.arange(vz)[None, :]).ravel()
            rows = rows[rows < nzz]
            count[x0:x1, rows] += 1
        inner = count[4:nxx - 4, 4:nzz - 4]
        assert inner.min() == 1 and inner.max() == 1, (nzz, nxx, vz, xw)
        # owned rows past the updated ones are rim rows (stored as zeros)
        assert count[:4].sum() == 0 and count[nxx - 4:].sum() == 0


@pytest.mark.parametrize("nzz,nxx", SHAPES)
def test_every_access_is_inside_its_allocation(nzz, nxx):
    ldz, fs = layout(nzz, nxx)
    for vz, xw, (reverse, steps), slots in itertools.product(
            (1, 2, 4), (3, 8, 81, 144, 1000), ((False, 3), (True, 3), (False, 4), (True, 4)),
            (1, 3)):
        if reverse and vz == 4:
            continue                          # no three- or four-step reverse kernel at VZ 4
        for buf, lo, hi in accesses(nzz, nxx, vz, xw, reverse, steps=steps):
            for slot in (0, slots - 1):
                size = slots * fs + GUARD
                # a one-slot coefficient array (sponge grid) is the tight case
                base = 0 if buf == "c" else slot * fs
                csize = fs + GUARD if buf == "c" else size
                assert base + lo >= 0, (buf, nzz, nxx, vz, xw, lo)
                assert base + hi < csize, (buf, nzz, nxx, vz, xw, hi, csize)


def test_the_model_catches_an_unclamped_halo():
    """The same model with the halo column clamped to [0, nxx) instead reads
    before the first slot: the out-of-bounds access the clamp prevents."""
    lows = [lo for buf, lo, hi in accesses(163, 283, 2, 81, False, halo_lo=0)]
    assert min(lows) == -8
    assert min(lo for buf, lo, hi in accesses(163, 283, 2, 81, False)) >= 0
