// layout.h -- device buffer sizes of an engine context and the extents its
// passes fill, as plain functions of the configuration.  Host-only (no HIP):
// the engine sizes its allocations and its fills from these, and a host test
// checks every fill extent against the allocation it writes.
#pragma once
#include "fwi2d.h"

#include <algorithm>

namespace fwi {

constexpr long kGuard = 1024;   // zeroed floats after each buffer a step kernel reads

// One padded grid as stored: columns of ldz = round_up(nzz, 4) floats, slots
// of fstride = round_up(nxx * ldz, 64) floats (256 B aligned).
struct GridDims {
    bool on;
    int nb, nzz, nxx, ldz;
    long fstride;
};

inline GridDims grid_dims(int nz, int nx, int nb)
{
    GridDims d{};
    d.on = nb > 0;
    if (!d.on) return d;
    d.nb = nb;
    d.nzz = nz + 2 * nb;
    d.nxx = nx + 2 * nb;
    d.ldz = (d.nzz + 3) & ~3;
    d.fstride = ((long)d.ldz * d.nxx + 63) & ~63L;
    return d;
}

// Buffer sizes (floats) of a context and the extents (floats) each pass
// fills.  grid[0] is the sponge grid, grid[1] the random-boundary grid.
struct Layout {
    GridDims grid[2];
    long slots;          // max_batch
    long fstride_max;    // the larger of the two grids' slot strides

    // allocations
    long wavefield() const { return slots * fstride_max + kGuard; }     // eight field buffers
    long rb_coef() const { return slots * grid[1].fstride + kGuard; }   // per-slot c, rbc grid
    long slot_acc() const { return slots * grid[1].fstride + kGuard; }  // per-slot image / illum
    long abc_coef() const { return grid[0].fstride + kGuard; }          // shared c, abc grid
    long shot_sum() const { return grid[1].fstride; }                   // image, gradient, illum

    // fills: the four fields of a pass on grid `kind`; the eight fields and
    // the per-slot accumulators around the autotuner's launches (which run
    // on both grids)
    long pass_fill(int kind, int nslots) const { return nslots * grid[kind].fstride; }
    long tune_fill(int nslots) const { return nslots * fstride_max; }
    long acc_fill(int nslots) const { return nslots * grid[1].fstride; }
};

inline Layout make_layout(const fwi_config &cfg)
{
    Layout L{};
    L.grid[0] = grid_dims(cfg.nz, cfg.nx, cfg.abc_nb);
    L.grid[1] = grid_dims(cfg.nz, cfg.nx, cfg.rbc_nb);
    L.slots = cfg.max_batch;
    L.fstride_max = std::max(L.grid[0].on ? L.grid[0].fstride : 0L,
                             L.grid[1].on ? L.grid[1].fstride : 0L);
    return L;
}

}  // namespace fwi
