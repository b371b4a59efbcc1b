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
columns >= 1 (clamped to >= 4 here).

The kernels on 64 x 64-cell LDS tiles (`fwdt_kernel`, `revt_kernel`, the
same tiling and step regions) are modelled at the end: their stores cover
the updated cells once, every LDS read of every step is a cell the load
phase filled or the previous step wrote, and the NW waves of a tile (thread
columns xg, xg + NW, ...) update each tile column once."""
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


# ---- forward kernel on 64 x 64-cell LDS tiles (fwdt_kernel)

TILE = 64


def tile_count(n, steps):
    own = TILE - 8 * steps
    return max(1, (n - 8 + own - 1) // own)


def tile_flow(nzz, nxx, steps):
    """Per tile: the LDS cells each step reads (current field at the stencil's
    17 points, older field and c at the cell), the cells it writes, and the
    grid cells the launch stores, in the kernel's tile coordinates."""
    h, own = 4 * steps, TILE - 8 * steps
    for tz in range(tile_count(nzz, steps)):
        for tx in range(tile_count(nxx, steps)):
            z0, x0 = 4 + tz * own - h, 4 + tx * own - h
            zlo, xlo = 4 + tz * own, 4 + tx * own
            zhi, xhi = min(zlo + own, nzz), min(xlo + own, nxx - 4)
            yield z0, x0, (zlo, zhi, xlo, xhi)


@pytest.mark.parametrize("nzz,nxx", [(321, 601), (81, 93), (601, 1301), (17, 19), (72, 72)])
@pytest.mark.parametrize("steps", [1, 2, 3, 4])
def test_tile_kernel_stores_cover_the_grid_once(nzz, nxx, steps):
    ldz, fs = layout(nzz, nxx)
    hits = np.zeros((nxx, ldz), np.int32)
    for z0, x0, (zlo, zhi, xlo, xhi) in tile_flow(nzz, nxx, steps):
        h = 4 * steps
        # the stored cells lie in the last step's region [h, 64 - h) of the tile
        assert zlo - z0 >= h and zhi - z0 <= TILE - h and xlo - x0 >= h and xhi - x0 <= TILE - h
        hits[xlo:xhi, zlo:zhi] += 1
        assert 0 <= xlo and xhi <= nxx - 4 and 4 <= zlo and zhi <= nzz
    assert (hits[4:nxx - 4, 4:nzz - 4] == 1).all()
    assert hits.max() == 1
    assert hits[:4].sum() == 0 and hits[nxx - 4:].sum() == 0 and hits[:, :4].sum() == 0
    assert hits[:, nzz:].sum() == 0


@pytest.mark.parametrize("steps", [1, 2, 3, 4])
def test_tile_kernel_reads_only_what_it_loaded_or_wrote(steps):
    """Each step k updates tile cells [4k, 64 - 4k)^2 from the current field
    at the stencil points and the older field at the cell; every such read is
    a cell the load phase filled (u_t, c everywhere; u_{t-1} on [4, 60)^2) or
    the previous step wrote, and stays inside the 64 x 64 tile."""
    n = TILE
    filled = [np.ones((n, n), bool), np.zeros((n, n), bool)]     # u[0] = u_t, u[1] = u_{t-1}
    filled[1][4:n - 4, 4:n - 4] = True
    offs = [(d, 0) for d in range(-4, 5)] + [(0, d) for d in range(-4, 5) if d]
    for k in range(1, steps + 1):
        cur, new = filled[(k + 1) & 1], filled[k & 1]
        lo, hi = 4 * k, n - 4 * k
        for x in range(lo, hi):
            for z in range(lo, hi):
                for dx, dz in offs:
                    assert 0 <= x + dx < n and 0 <= z + dz < n
                    assert cur[x + dx, z + dz], (k, x, z, dx, dz)
                assert new[x, z], (k, x, z)          # the older field at the cell
        region = np.zeros((n, n), bool)
        region[lo:hi, lo:hi] = True
        filled[k & 1] = new | region
        # one step: u_{t-1} is read only on the tile's own cells, which its
        # result overwrites in place in global memory
        if steps == 1:
            h = 4
            assert lo == h and hi == n - h


@pytest.mark.parametrize("nw", [4, 8, 16])
def test_tile_threads_cover_each_column_once(nw):
    cols = [xg + nw * m for xg in range(nw) for m in range(TILE // nw)]
    assert sorted(cols) == list(range(TILE))
