// kernels.h -- launch interface of the HIP kernels (device code lives in
// propagate.hip).  Host-only declarations: plain structs and launchers.
#pragma once
#include "rbc.h"
#include <hip/hip_runtime.h>
#include <cstdint>

namespace fwi {

// 8th-order second-derivative weights (src/modeling/modeling.cuh:9-13).
constexpr float kC1 = 0.001785714285f;
constexpr float kC2 = 0.025396825396f;
constexpr float kC3 = 0.2f;
constexpr float kC4 = 1.6f;

// Padded grid as stored on the device: column j of shot slot s starts at
// s*fstride + j*ldz; rows nzz..ldz-1 are zero padding.
struct Grid {
    int nzz, nxx, ldz, nb;
    long fstride;      // floats per field per shot slot
};

// Wave tiling of one launch: each 64-lane wave computes 64*vz rows of one
// x-chunk of xw columns of one shot slot and writes the inner
// 64*vz - 8*(steps-1) of them (steps per launch: 1-4).
// vz = kTileVz selects the kernels on 64 x 64-cell LDS tiles instead
// (nzseg x nxchunk tiles of (64 - 8 steps)^2 own cells; xw = the tile columns
// each wave updates, 16, 8 or 4: 4, 8 or 16 waves per tile).
constexpr int kTileVz = 16;
constexpr int kTile = 64;          // tile edge in cells
struct Tiling {
    int vz;            // rows per lane: 1, 2 or 4; or kTileVz
    int xw;            // columns per wave
    int nzseg, nxchunk;
    bool staged;       // every wave's window fits the LDS tables below
};

// Per-wave LDS copies of the receiver table (and sponge x factors) cover a
// window of at most kStageCols columns holding at most kStageRec receivers;
// a launch of s steps has the window [x0 - halo, x1 + halo), halo = 4(s-1).
constexpr int kStageCols = 160;
constexpr int kStageRec = 256;
// the four-step reverse kernel's table (its LDS also holds 13 columns of c and
// 9 of image partial sums per wave)
constexpr int kStageRec4 = 176;
inline int window_halo(int steps) { return 4 * (steps - 1); }

// Receivers of one grid, sorted by column (CSR over columns).
struct RecTable {
    const int *col_ptr;   // [nxx+1]
    const int *iz;        // [nrec] row of each sorted receiver
    const int *id;        // [nrec] original receiver index
    int nrec;
};

// One forward launch at step t.  Single step: reads cur = u_t and
// old = u_{t-1}, writes u_{t+1} over old (out1 == old).  s = 2, 3 or 4 steps:
// reads cur, old; writes u_{t+s-1} to out1 and u_{t+s} to out2 (four
// distinct buffers).
struct FwdStep {
    Grid g;
    const float *cur;
    const float *old;
    float *out1, *out2;
    const float *c;            // dt^2 v^2 / dh^2
    long cstride;              // per-slot stride of c (0 = shared by all slots)
    const float *b1d;          // sponge table [nb] (abc only)
    int abc;
    const int *src_iz, *src_ix;   // per global shot (padded grid)
    const int *shot0;          // device scalar: first global shot of the batch
    const float *wav;          // idh2 * w[t], t < nt
    int nt, tlag, T;
    RecTable rec;
    float *seis;               // [slot][nrec][nt] or nullptr
};

// One reverse launch at reverse step k.  S: cur = u_{T-1-k}, old = u_{T-k};
// R: cur = R_k, old = R_{k-1}.  Single step writes over the olds; a launch
// of s = 2, 3 or 4 steps writes steps k+s-1 / k+s to the out buffers.
struct RevStep {
    Grid g;
    const float *scur, *sold;
    float *sout1, *sout2;
    const float *rcur, *rold;
    float *rout1, *rout2;
    const float *c;
    long cstride;
    const int *shot0;
    const float *data;         // injected data [shot][nrec][nt]
    int data_by_shot;          // 1: row = shot0+slot (observed), 0: row = slot
    float idh2;
    int nt, T;
    RecTable rec;
    float *acc, *ill;          // per-slot accumulators [slot][field]
    int mode;                  // 0 image, 1 gradient
};

// steps = 1-4 time steps per launch (the three- and four-step reverse
// kernels: vz 1 or 2)
hipError_t launch_fwd_step(const FwdStep &a, const Tiling &tl, int nslots, int t, int steps,
                           hipStream_t s);
hipError_t launch_rev_step(const RevStep &a, const Tiling &tl, int nslots, int k, int steps,
                           hipStream_t s);
// z segments of a launch shape (a launch of s steps writes 64*vz - 8(s-1) rows per wave)
// and its x chunks; for kTileVz the tile rows and columns
int zsegments(int nzz, int vz, int steps);
int xchunks(int nxx, int vz, int xw, int steps);
// Waves of the random-boundary forward (kind 0), imaging reverse (kind 1) or
// sponge forward (kind 2) step kernel that the current device holds at once
// (0 if unknown).
int resident_waves(int kind, int steps, int vz);

// Adds idh2*w[0] at each slot's source cell of `field` (before step 0).
hipError_t launch_fwd_init(const FwdStep &a, float *field, int nslots, hipStream_t s);
// Adds idh2*data[nt-1] at the receivers of `field` (before reverse step 0).
hipError_t launch_rev_init(const RevStep &a, float *field, int nslots, hipStream_t s);

// Fill of n floats (graph-safe replacement for memset).
hipError_t launch_fill(float *p, long n, float v, hipStream_t s);
hipError_t launch_set_int(int *p, int v, hipStream_t s);

// Edge-replication pad of vp[nx][nz] into out (grid g, one slot).
hipError_t launch_pad_model(const float *vp, int nz, int nx, const Grid &g, float *out,
                            hipStream_t s);
// c = k * v^2 over one slot.
hipError_t launch_coef(const float *v, const Grid &g, float k, float *c, hipStream_t s);
// Random-boundary velocity of nslots slots (global shots shot0+slot) from the
// padded base vpad; writes c (if c) and v (if v_out) per slot.
hipError_t launch_rb(const float *vpad, const Grid &g, int nslots, const int *shot0,
                     uint32_t iteration, uint64_t seed, float amp_abs, float amp_rel,
                     float vlo, float vhi, float k, float *c, float *v_out, hipStream_t s);
// The reference-style random boundary (Gaussian ramps and the blob table of
// rbc.h, one table segment per slot) of nslots slots; like launch_rb.
hipError_t launch_rbg(const float *vpad, const Grid &g, int nslots, const int *shot0,
                      uint32_t iteration, uint64_t seed, float sigma, float vlo, float vhi,
                      float k, float dh, const RbBlob *blob, const int32_t *bstart, int nbx,
                      int nbz, float bin, float *c, float *v_out, hipStream_t s);
// residual = obs[shot0+slot] - cal (in place in cal), FP64 partial sums.
hipError_t launch_residual(const float *obs, float *cal, const int *shot0, long per_shot,
                           int nslots, double *partials, int nblocks, hipStream_t s);
hipError_t launch_sum_partials(const double *partials, int n, double *out, hipStream_t s);
// dst[cell] += sum_slot src[slot][cell] over nslots, then src = 0.
hipError_t launch_reduce_slots(float *src, long fstride, long ncell, int nslots, float *dst,
                               hipStream_t s);
// Crop of a padded field into [nx][nz].
hipError_t launch_crop(const float *src, const Grid &g, int nz, int nx, float *dst,
                       hipStream_t s);
// dst = src over n floats (n a multiple of 4), float4 per lane.
hipError_t launch_copy(const float *src, float *dst, long n, hipStream_t s);
// Sets seis[slot][r][0] = 0 for all slots.
hipError_t launch_zero_sample0(float *seis, int nrec, int nt, int nslots, hipStream_t s);

}  // namespace fwi
