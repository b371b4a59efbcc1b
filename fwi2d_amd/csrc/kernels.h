// kernels.h -- launch interface of the HIP kernels (device code lives in
// propagate.hip).  Host-only declarations: plain structs and launchers.
#pragma once
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

// Wave tiling of one time step: each 64-lane wave owns 64*vz rows of one
// x-chunk of xw columns of one shot slot.
struct Tiling {
    int vz;            // rows per lane: 1, 2 or 4
    int xw;            // columns per wave
    int nzseg, nxchunk;
};

// Receivers of one grid, sorted by column (CSR over columns).
struct RecTable {
    const int *col_ptr;   // [nxx+1]
    const int *iz;        // [nrec] row of each sorted receiver
    const int *id;        // [nrec] original receiver index
    int nrec;
};

struct FwdStep {
    Grid g;
    float *fa, *fb;            // field buffers; step t reads cur = (t&1 ? fb : fa)
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

struct RevStep {
    Grid g;
    float *sa, *sb;            // source wavefield pair; step k: cur = (k&1 ? sa : sb)
    float *ra, *rb;            // receiver wavefield pair; step k: cur = (k&1 ? rb : ra)
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

hipError_t launch_fwd_step(const FwdStep &a, const Tiling &tl, int nslots, int t,
                           hipStream_t s);
hipError_t launch_rev_step(const RevStep &a, const Tiling &tl, int nslots, int k,
                           hipStream_t s);

// Adds idh2*w[0] at each slot's source cell of field fa (before step 0).
hipError_t launch_fwd_init(const FwdStep &a, int nslots, hipStream_t s);
// Adds idh2*data[nt-1] at the receivers of field ra (before reverse step 0).
hipError_t launch_rev_init(const RevStep &a, int nslots, hipStream_t s);

// Zero-fill of n floats (graph-safe replacement for memset).
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
// Sets seis[slot][r][0] = 0 for all slots.
hipError_t launch_zero_sample0(float *seis, int nrec, int nt, int nslots, hipStream_t s);

}  // namespace fwi
