// propagate.hip -- CDNA4 (gfx950) kernels of the 2-D acoustic engine.
//
// Time stepping is the reference's 8th-order-space / 2nd-order-time
// leapfrog (src/modeling/modeling.cu:301-345, src/migration/migration.cu:131-180,
// src/inversion/inversion.cu:395-444), re-laid out for the hardware:
//
//  * z (the contiguous axis of every reference array) maps to lanes: a
//    64-lane wave owns 64*VZ consecutive rows, VZ = 1, 2 or 4 per lane, so
//    each column load is one coalesced 256 B .. 1 KiB access per wave;
//  * the wave marches along x over xw columns and keeps the 9 columns the
//    x-derivative needs in a register queue, so the current field is read
//    from memory about once per step (plus an 8-column warm-up per chunk);
//  * the +-4 z-neighbours come from the neighbouring lanes by DPP whole-wave
//    shifts; the 4 rows beyond each end of the wave come from one prefetched
//    float4 "halo" load per column (one- and two-step kernels), or from one
//    float per lane loaded beside the column's main load (mostly the same
//    cache lines) and kept with it in the queue (three- and four-step);
//  * one to four time steps per launch (temporal blocking): each step
//    after the first is computed on 8 fewer rows and columns than the one
//    before, from a register queue of the previous step's values; c, read
//    for the first step, waits in an LDS ring for the later ones; a launch
//    writes only the two fields the next one needs, so the HBM bytes of a
//    launch do not grow with its steps;
//  * the three- and four-step kernels unroll their column loop over the
//    nine phases of the queues, so the queues advance by renaming, not by
//    moves, and their loads are branch-free (columns clamped into the grid);
//  * source injection, the Cerjan sponge, receiver sampling (forward) and
//    data injection, imaging / gradient and illumination (reverse) are fused
//    into the same pass; the receiver tables and data of a wave's columns
//    are staged in LDS once per launch (a second build of each kernel reads
//    them from global memory, for launches whose windows do not fit);
//  * blocks are renumbered so that each XCD receives a contiguous run of
//    tiles, and in the three- and four-step kernels odd x-chunks march
//    right to left: two neighbouring chunks then reach the columns they
//    share at the same time, which keeps the second read of those halo
//    columns in the caches (the stencil is symmetric and each sum is
//    commutative, so the mirrored march gives the same bits);
//  * for one or a few shots, where the column marches are too few to fill
//    the device and each one's length sets the launch time, the forward and
//    reverse steps also run on 64 x 64-cell tiles held in LDS (fwdt_kernel,
//    revt_kernel): one workgroup per tile, the z neighbours by DPP shifts,
//    the x neighbours from LDS, one to four steps per launch.
//
// FP32 operation order: the Laplacian's nearest ring is summed as
// differences from the centre cell (the farther rings as sums minus 4u) and
// the update as (p + (p - old)) + c*lap, with c = dt^2 v^2/dh^2.  This is the
// reference's recurrence with less rounding error (about 4-5x less than the
// reference's order against an FP64 evaluation of the same recurrence, over
// 2,000+ steps).  Every multiply-add that feeds a stored value is an explicit
// fmaf, so all launch shapes (one to four steps per launch, any tiling)
// produce bit-identical fields, images and gradients.
//
// Source injection is folded into the write of the field it enters:
// the field written at step t receives idh2*w[t+1], which is what the
// reference's step t+1 adds before reading it (modeling.cu:316-317).  With
// the sponge on, the reference damps both the new and the current field in
// place after each step (modeling.cu:337-343); here the stored field carries
// the first factor (as the new field) and the second one is applied when it
// is read back as the previous field one step later:
//   u_{t+1} = d * ((u_t + (u_t - d*u_{t-1})) + c*lap u_t)  [+ idh2 w[t+1]].
#include "kernels.h"

namespace fwi {

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

template <int VZ>
__device__ __forceinline__ void vload(float (&r)[VZ], const float *p)
{
    if constexpr (VZ == 4) {
        const float4 v = *reinterpret_cast<const float4 *>(p);
        r[0] = v.x; r[1] = v.y; r[2] = v.z; r[3] = v.w;
    } else if constexpr (VZ == 2) {
        const float2 v = *reinterpret_cast<const float2 *>(p);
        r[0] = v.x; r[1] = v.y;
    } else {
        r[0] = *p;
    }
}

template <int VZ>
__device__ __forceinline__ void vstore(float *p, const float (&r)[VZ])
{
    if constexpr (VZ == 4) {
        *reinterpret_cast<float4 *>(p) = make_float4(r[0], r[1], r[2], r[3]);
    } else if constexpr (VZ == 2) {
        *reinterpret_cast<float2 *>(p) = make_float2(r[0], r[1]);
    } else {
        *p = r[0];
    }
}

template <int VZ>
__device__ __forceinline__ void vzero(float (&r)[VZ])
{
#pragma unroll
    for (int k = 0; k < VZ; ++k) r[k] = 0.f;
}

// Compile-time queue phase of an unrolled column loop.
template <int N>
struct Phase {
    static constexpr int value = N;
};

// Whole-wave lane shifts (GFX9 DPP wave_shr:1 / wave_shl:1): lane i takes
// lane i-1's (shr) or lane i+1's (shl) value; the lane with no source (0 or
// 63) takes `edge`.  Plain VALU moves, no LDS traffic.
__device__ __forceinline__ float wave_shr1(float src, float edge)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(src),
                                                      0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float wave_shl1(float src, float edge)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(src),
                                                      0x130, 0xF, 0xF, false));
}

// The 8 rows around a wave's 64*VZ rows [z0, z0 + 64*VZ) of one column:
// lanes 0-31 get rows z0-4 .. z0-1, lanes 32-63 rows z0+64*VZ .. +3 (one
// aligned float4 per lane; z0 is a multiple of 4).  Zero outside [lo, hi).
template <int VZ>
__device__ __forceinline__ float4 ldhalo(const float *base, int col, int lo, int hi, int ldz,
                                         int z0, int lane)
{
    if (col >= lo && col < hi)
        return *reinterpret_cast<const float4 *>(base + (long)col * ldz +
                                                 (lane < 32 ? z0 - 4 : z0 + 64 * VZ));
    return make_float4(0.f, 0.f, 0.f, 0.f);
}

// ext[0 .. VZ+7] = rows zb-4 .. zb+VZ+3 of one column: the lane's own VZ
// rows and 4 rows above and below, shifted in from the neighbouring lanes;
// lane 0 and lane 63 take the rows beyond the wave from the halo vector h.
template <int VZ>
__device__ __forceinline__ void zext(float (&ext)[VZ + 8], const float (&own)[VZ], float4 h)
{
#pragma unroll
    for (int r = 0; r < VZ; ++r) ext[4 + r] = own[r];
    if constexpr (VZ == 4) {
        ext[0] = wave_shr1(own[0], h.x);
        ext[1] = wave_shr1(own[1], h.y);
        ext[2] = wave_shr1(own[2], h.z);
        ext[3] = wave_shr1(own[3], h.w);
        ext[8] = wave_shl1(own[0], h.x);
        ext[9] = wave_shl1(own[1], h.y);
        ext[10] = wave_shl1(own[2], h.z);
        ext[11] = wave_shl1(own[3], h.w);
    } else if constexpr (VZ == 2) {
        ext[2] = wave_shr1(own[0], h.z);
        ext[3] = wave_shr1(own[1], h.w);
        ext[0] = wave_shr1(ext[2], h.x);
        ext[1] = wave_shr1(ext[3], h.y);
        ext[6] = wave_shl1(own[0], h.x);
        ext[7] = wave_shl1(own[1], h.y);
        ext[8] = wave_shl1(ext[6], h.z);
        ext[9] = wave_shl1(ext[7], h.w);
    } else {
        ext[3] = wave_shr1(own[0], h.w);
        ext[2] = wave_shr1(ext[3], h.z);
        ext[1] = wave_shr1(ext[2], h.y);
        ext[0] = wave_shr1(ext[1], h.x);
        ext[5] = wave_shl1(own[0], h.x);
        ext[6] = wave_shl1(ext[5], h.y);
        ext[7] = wave_shl1(ext[6], h.z);
        ext[8] = wave_shl1(ext[7], h.w);
    }
}


// Tile of a 1-D launch after the XCD renumbering: each of the 8 XCDs gets a
// contiguous run of per = ceil(ntiles/8) logical tiles.  Returns -1 for the
// blocks past the last tile.
__device__ __forceinline__ int xcd_tile(int ntiles)
{
    const int id = blockIdx.x;
    const int per = (ntiles + 7) >> 3;
    const int logical = (id & 7) * per + (id >> 3);
    return logical < ntiles ? logical : -1;
}

// Laplacian (times dh^2) from a 9-column register queue whose column j-4+m
// sits in slot (m + OFF) % 9, and the z extension of the centre column.  The
// four nearest neighbours (weight 1.6) enter as differences from the centre,
// which keeps the FP32 cancellation against the centre term small; the
// farther rings (weights 0.2 and below) as sums minus 4u.  The weight sum is
// one explicit FMA chain so that every kernel rounds it identically.
template <int VZ, int OFF>
__device__ __forceinline__ float lap8r(const float (&q)[9][VZ], const float (&ext)[VZ + 8],
                                       int r)
{
    constexpr int P0 = (0 + OFF) % 9, P1 = (1 + OFF) % 9, P2 = (2 + OFF) % 9,
                  P3 = (3 + OFF) % 9, P4 = (4 + OFF) % 9, P5 = (5 + OFF) % 9,
                  P6 = (6 + OFF) % 9, P7 = (7 + OFF) % 9, P8 = (8 + OFF) % 9;
    const float p = q[P4][r];
    const float s1 = ((q[P3][r] - p) + (q[P5][r] - p)) + ((ext[r + 3] - p) + (ext[r + 5] - p));
    const float s2 = fmaf(-4.f, p, (q[P2][r] + q[P6][r]) + (ext[r + 2] + ext[r + 6]));
    const float s3 = fmaf(-4.f, p, (q[P1][r] + q[P7][r]) + (ext[r + 1] + ext[r + 7]));
    const float s4 = fmaf(-4.f, p, (q[P0][r] + q[P8][r]) + (ext[r + 0] + ext[r + 8]));
    return fmaf(kC4, s1, fmaf(-kC3, s2, fmaf(kC2, s3, -kC1 * s4)));
}

// lap8r with the z-part of the nearest ring taken from the row differences
// dz[m] = ext[m + 4] - ext[m + 3] (zdiff), which neighbouring rows of a lane
// share: (ext[r+3] - p) + (ext[r+5] - p) is -dz[r] + dz[r+1], the same
// rounded values summed in the same order, so the result is bit-identical
// to lap8r with VZ + 1 subtractions per lane instead of 2 VZ.
template <int VZ, int OFF>
__device__ __forceinline__ float lap8z(const float (&q)[9][VZ], const float (&ext)[VZ + 8],
                                       const float (&dz)[VZ + 1], int r)
{
    constexpr int P0 = (0 + OFF) % 9, P1 = (1 + OFF) % 9, P2 = (2 + OFF) % 9,
                  P3 = (3 + OFF) % 9, P4 = (4 + OFF) % 9, P5 = (5 + OFF) % 9,
                  P6 = (6 + OFF) % 9, P7 = (7 + OFF) % 9, P8 = (8 + OFF) % 9;
    const float p = q[P4][r];
    const float s1 = ((q[P3][r] - p) + (q[P5][r] - p)) + (dz[r + 1] - dz[r]);
    const float s2 = fmaf(-4.f, p, (q[P2][r] + q[P6][r]) + (ext[r + 2] + ext[r + 6]));
    const float s3 = fmaf(-4.f, p, (q[P1][r] + q[P7][r]) + (ext[r + 1] + ext[r + 7]));
    const float s4 = fmaf(-4.f, p, (q[P0][r] + q[P8][r]) + (ext[r + 0] + ext[r + 8]));
    return fmaf(kC4, s1, fmaf(-kC3, s2, fmaf(kC2, s3, -kC1 * s4)));
}
template <int VZ>
__device__ __forceinline__ void zdiff(float (&dz)[VZ + 1], const float (&ext)[VZ + 8])
{
#pragma unroll
    for (int m = 0; m <= VZ; ++m) dz[m] = ext[m + 4] - ext[m + 3];
}

// acc + a*b with the product rounded on its own (no FMA): the one-step and
// two-step kernels then accumulate images and illumination identically.
__device__ __forceinline__ float mul_nc(float a, float b)
{
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ float add_nc(float a, float b)
{
#pragma clang fp contract(off)
    return a + b;
}

// Sponge factors split by axis: dz(i) = b1d of the row's edge distance in
// the z pads (1 elsewhere), dx(j) likewise; the reference's lookup
// (modeling.cu:355-402) is d = 1 (interior), dz (x inside), dx (z inside),
// dz + dx - 1 (corners).  All four are d = dx + (dz - 1): dz - 1 is exact
// for dz >= 1/2 (Sterbenz) and 1 + (dz - 1) then rounds back to dz, so the
// edges and the interior come out exactly and the corners are the sum.  One
// add per cell and no per-lane or per-column flags.
__device__ __forceinline__ float pad_factor(const float *b1d, int k, int n, int nb, bool &in)
{
    k = min(max(k, 0), n - 1);
    in = (k >= nb) && (k < n - nb);
    return in ? 1.f : b1d[(k < nb) ? k : (n - 1 - k)];
}
__device__ __forceinline__ float sponge(float dx, float dzm1)
{
    return dx + dzm1;
}
// dz - 1 of a lane's VZ rows.
template <int VZ>
__device__ __forceinline__ void row_factors(const float *b1d, int zb, const Grid &g,
                                            float (&dzm1)[VZ])
{
#pragma unroll
    for (int r = 0; r < VZ; ++r) {
        bool in;
        dzm1[r] = pad_factor(b1d, zb + r, g.nzz, g.nb, in) - 1.f;
    }
}

// Cerjan factor of cell (i,j): 1 inside, b1d on the edges, b1d[a]+b1d[b]-1
// in the corners (modeling.cu:355-402), in the form of sponge() below.
// i is clamped into the grid.
__device__ __forceinline__ float damper(const float *b1d, int i, int j, const Grid &g)
{
    bool in;
    const float dz = pad_factor(b1d, i, g.nzz, g.nb, in);
    const float dx = pad_factor(b1d, j, g.nxx, g.nb, in);
    return sponge(dx, dz - 1.f);
}

// One leapfrog update with explicit FMAs:
//   no sponge: u_new = (p + (p - o)) + c lap
//   sponge:    u_new = d ((p + (p - d o)) + c lap)
template <bool ABC>
__device__ __forceinline__ float leapfrog(float p, float o, float c, float lap, float d)
{
    if constexpr (ABC) return d * fmaf(c, lap, p + fmaf(-d, o, p));
    else return fmaf(c, lap, p + (p - o));
}

// Adds the source sample w to the source cell (row siz of column six) when
// column j holds it; the cell is inside the rim, so adding after the rim
// select gives the same value as adding before it.
template <int VZ>
__device__ __forceinline__ void add_source(float (&v)[VZ], int j, int six, int zb, int siz,
                                           float w)
{
    if (j == six) {
#pragma unroll
        for (int r = 0; r < VZ; ++r)
            if (zb + r == siz) v[r] += w;
    }
}

// z-extension of a field that exists only in registers (the inner field of a
// two-step launch): rows beyond the wave are never used for a written cell,
// so the edge lanes take zeros (DPP bound_ctrl).
__device__ __forceinline__ float wave_shr1z(float src)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(src), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float wave_shl1z(float src)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(src), 0x130, 0xF, 0xF, true));
}

template <int VZ>
__device__ __forceinline__ void zext_local(float (&ext)[VZ + 8], const float (&own)[VZ])
{
#pragma unroll
    for (int r = 0; r < VZ; ++r) ext[4 + r] = own[r];
    if constexpr (VZ == 4) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ext[k] = wave_shr1z(own[k]);
            ext[8 + k] = wave_shl1z(own[k]);
        }
    } else if constexpr (VZ == 2) {
        ext[2] = wave_shr1z(own[0]);
        ext[3] = wave_shr1z(own[1]);
        ext[0] = wave_shr1z(ext[2]);
        ext[1] = wave_shr1z(ext[3]);
        ext[6] = wave_shl1z(own[0]);
        ext[7] = wave_shl1z(own[1]);
        ext[8] = wave_shl1z(ext[6]);
        ext[9] = wave_shl1z(ext[7]);
    } else {
        ext[3] = wave_shr1z(own[0]);
        ext[2] = wave_shr1z(ext[3]);
        ext[1] = wave_shr1z(ext[2]);
        ext[0] = wave_shr1z(ext[1]);
        ext[5] = wave_shl1z(own[0]);
        ext[6] = wave_shl1z(ext[5]);
        ext[7] = wave_shl1z(ext[6]);
        ext[8] = wave_shl1z(ext[7]);
    }
}

// Column `col` of a slot field into registers; zero outside the grid's
// columns.  The load is not masked per lane: rows at or past ldz read the
// next column's first rows (zero rim rows of a wavefield) or the zeroed
// guard after each buffer, and only feed rim rows, which are never written.
template <int VZ>
__device__ __forceinline__ void ldcol(float (&r)[VZ], const float *base, int col, int nxx,
                                      int ldz, int zb)
{
    if (col >= 0 && col < nxx) vload<VZ>(r, base + (long)col * ldz + zb);
    else vzero<VZ>(r);
}

// Loads for the unrolled three- and four-step kernels, with no branch: the column is
// clamped into the grid, so a column beyond the edge reads the edge column,
// which is a zero rim column of every wavefield (sources and receivers sit
// inside the rim).  Such values only feed rim columns.
// The address is a wave-uniform column pointer plus the lane's unsigned row
// offset, so the access takes its base in SGPRs and one offset VGPR.
template <int VZ>
__device__ __forceinline__ void ldcolc(float (&r)[VZ], const float *base, int col, int nxx,
                                       int ldz, unsigned zl)
{
    col = min(max(col, 0), nxx - 1);
    vload<VZ>(r, (base + (long)col * ldz) + zl);
}
// The halo rows of column col clamped into [4, nxx-4), the columns whose
// values it feeds.  The first row segment starts at z0 = -4, so its rows
// z0-4 .. z0-1 lie in the column before: column 4 keeps that read inside
// the slot's field (a column below 4 would not).
template <int VZ>
__device__ __forceinline__ float4 ldhaloc(const float *base, int col, int nxx, int ldz, int z0,
                                          int lane)
{
    col = min(max(col, 4), nxx - 5);
    const unsigned off = lane < 32 ? 0u : unsigned(64 * VZ + 4);
    return *reinterpret_cast<const float4 *>((base + (long)col * ldz + (z0 - 4)) + off);
}

// The eight halo rows of ldhaloc's column in one VGPR: lanes 0-3 hold rows
// z0-4 .. z0-1 and lanes 60-63 rows z0+64*VZ .. z0+64*VZ+3 (the lanes in
// between repeat them), the addresses ldhaloc's float4 covers.
template <int VZ>
__device__ __forceinline__ float ldhalo1(const float *base, int col, int nxx, int ldz, int z0,
                                         int lane)
{
    col = min(max(col, 4), nxx - 5);
    const unsigned off = (lane < 32 ? 0u : unsigned(64 * VZ + 4)) + unsigned(lane & 3);
    return (base + (long)col * ldz + (z0 - 4))[off];
}

// Lane 4q + K's value in every lane of quad q (DPP quad_perm [K,K,K,K]): for
// a halo of ldhalo1, lane 0 receives row K above the wave and lane 63 row K
// below it, the two lanes whose edge values zext reads.
template <int K>
__device__ __forceinline__ float quad_bcast(float h)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(h), K * 0x55, 0xF, 0xF, false));
}

// zext from a halo of ldhalo1.
template <int VZ>
__device__ __forceinline__ void zext1(float (&ext)[VZ + 8], const float (&own)[VZ], float hv)
{
    zext<VZ>(ext, own,
             make_float4(quad_bcast<0>(hv), quad_bcast<1>(hv), quad_bcast<2>(hv),
                         quad_bcast<3>(hv)));
}

// Receivers in column j whose row is in this lane: seis[rec][tsamp] = d * v[row].
template <int VZ, bool ABC>
__device__ __forceinline__ void sample_receivers(const FwdStep &a, const float (&v)[VZ], int j,
                                                 int zb, float *seis, int tsamp)
{
    const int k0 = a.rec.col_ptr[j], k1 = a.rec.col_ptr[j + 1];
    for (int k = k0; k < k1; ++k) {
        const int rr = a.rec.iz[k] - zb;
#pragma unroll
        for (int r = 0; r < VZ; ++r) {
            if (rr == r) {
                float val = v[r];
                if constexpr (ABC) val *= damper(a.b1d, zb + r, j, a.g);
                seis[(long)a.rec.id[k] * a.nt + tsamp] = val;
            }
        }
    }
}

// sample3_staged from the global receiver table.
template <int VZ, bool ABC>
__device__ __forceinline__ void sample3_receivers(const FwdStep &a, const float (&v1)[VZ],
                                                  const float (&v2)[VZ], const float (&v3)[VZ],
                                                  int j, int zb, float *seis, int tsamp, bool r1,
                                                  bool r2, bool r3)
{
    const int k0 = a.rec.col_ptr[j], k1 = a.rec.col_ptr[j + 1];
    for (int k = k0; k < k1; ++k) {
        const int rr = a.rec.iz[k] - zb;
#pragma unroll
        for (int r = 0; r < VZ; ++r)
            if (rr == r) {
                const float d = ABC ? damper(a.b1d, zb + r, j, a.g) : 1.f;
                const long base = (long)a.rec.id[k] * a.nt + tsamp;
                if (r1) seis[base] = v1[r] * d;
                if (r2) seis[base + 1] = v2[r] * d;
                if (r3) seis[base + 2] = v3[r] * d;
            }
    }
}

// Adds idh2 * data[rec][trow] to the receivers of column j in this lane.
template <int VZ>
__device__ __forceinline__ void inject_receivers(const RevStep &a, float (&v)[VZ], int j, int zb,
                                                 const float *dat, int trow)
{
    const int k0 = a.rec.col_ptr[j], k1 = a.rec.col_ptr[j + 1];
    for (int m = k0; m < k1; ++m) {
        const int rr = a.rec.iz[m] - zb;
        const float val = dat[(long)a.rec.id[m] * a.nt + trow];
#pragma unroll
        for (int r = 0; r < VZ; ++r) v[r] = fmaf(a.idh2, rr == r ? val : 0.f, v[r]);
    }
}

// Per-wave LDS copy of the receiver table (and, for the sponge, of the x
// factors) over the wave's column window [xa, xa + nc).  It is filled once
// at kernel start so that the column march never waits on a global load whose
// result it needs at once: such a wait also drains the loads prefetched for
// the next column.  Each step kernel is built twice: staged (ST), launched
// when every window of the launch fits the tables (the host checks it from
// the receiver columns), and with the global-memory path only.  One kernel
// holding both paths would join them inside the unrolled column loop, where
// the compiler then drains all outstanding loads at each join.

struct FwdStage {
    int cp[kStageCols + 1];     // col_ptr[xa + c] - col_ptr[xa]
    int iz[kStageRec];
    int id[kStageRec];
    float dx[kStageCols];       // sponge x factor of column xa + c
};
struct RevStage {
    int cp[kStageCols + 1];
    int iz[kStageRec];
    float v[3][kStageRec];      // data rows of the launch's steps 1-3 (rows in the wave)
};

// Column offsets and rows of the window's receivers; returns false (for the
// whole wave) when the window does not fit.
template <class Stage>
__device__ __forceinline__ bool stage_table(const RecTable &rt, int xa, int nc, int nxx,
                                            Stage &st, int lane, int &e0, int &ne)
{
    const int ca = min(max(xa, 0), nxx), cb = min(max(xa + nc, 0), nxx);
    e0 = rt.col_ptr[ca];
    ne = rt.col_ptr[cb] - e0;
    if (nc > kStageCols || ne > int(sizeof(st.iz) / sizeof(int))) return false;
    for (int c = lane; c <= nc; c += 64) st.cp[c] = rt.col_ptr[min(max(xa + c, 0), nxx)] - e0;
    for (int e = lane; e < ne; e += 64) st.iz[e] = rt.iz[e0 + e];
    return true;
}

__device__ __forceinline__ bool stage_fwd(const FwdStep &a, bool rec, bool abc, int xa, int nc,
                                          FwdStage &st, int lane)
{
    bool ok = nc <= kStageCols;
    if (abc && ok)
        for (int c = lane; c < nc; c += 64) {
            bool in;
            st.dx[c] = pad_factor(a.b1d, xa + c, a.g.nxx, a.g.nb, in);
        }
    if (rec && ok) {
        int e0, ne;
        ok = stage_table(a.rec, xa, nc, a.g.nxx, st, lane, e0, ne);
        if (ok) {
            for (int e = lane; e < ne; e += 64) st.id[e] = a.rec.id[e0 + e];
        }
    }
    __syncthreads();
    return ok;
}

// Rows [zlo, zhi) are the wave's; data of other rows is never read.
__device__ __forceinline__ bool stage_rev(const RevStep &a, const float *dat, int xa, int nc,
                                          int zlo, int zhi, bool on1, int trow1, bool on2,
                                          int trow2, bool on3, int trow3, RevStage &st, int lane)
{
    int e0 = 0, ne = 0;
    const bool ok = (on1 || on2 || on3) && stage_table(a.rec, xa, nc, a.g.nxx, st, lane, e0, ne);
    if (ok) {
        for (int e = lane; e < ne; e += 64) {
            const int iz = st.iz[e];        // written by this lane
            float v1 = 0.f, v2 = 0.f, v3 = 0.f;
            if (iz >= zlo && iz < zhi) {
                const float *row = dat + (long)a.rec.id[e0 + e] * a.nt;
                if (on1) v1 = row[trow1];
                if (on2) v2 = row[trow2];
                if (on3) v3 = row[trow3];
            }
            st.v[0][e] = v1;
            st.v[1][e] = v2;
            st.v[2][e] = v3;
        }
    }
    __syncthreads();
    return ok;
}

// The four-step reverse kernel's table: data rows of steps 1-4, at most
// kStageRec4 receivers (its LDS also holds 13 columns of c and 9 of image
// partial sums).
struct RevStage4 {
    int cp[kStageCols + 1];
    int iz[kStageRec4];
    float v[4][kStageRec4];
};

__device__ __forceinline__ bool stage_rev4(const RevStep &a, const float *dat, int xa, int nc,
                                           int zlo, int zhi, int k, RevStage4 &st, int lane)
{
    int e0 = 0, ne = 0;
    const bool ok = (k + 1 < a.nt) && stage_table(a.rec, xa, nc, a.g.nxx, st, lane, e0, ne);
    if (ok) {
        for (int e = lane; e < ne; e += 64) {
            const int iz = st.iz[e];        // written by this lane
            float v[4] = {0.f, 0.f, 0.f, 0.f};
            if (iz >= zlo && iz < zhi) {
                const float *row = dat + (long)a.rec.id[e0 + e] * a.nt;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (k + 1 + q < a.nt) v[q] = row[a.nt - 2 - k - q];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) st.v[q][e] = v[q];
        }
    }
    __syncthreads();
    return ok;
}

// seis[rec][tsamp] = d * v for the receivers of window column c in this
// lane, d the sponge factor of the cell (dx of the column, dz - 1 of the
// row) or 1.
template <int VZ, bool ABC>
__device__ __forceinline__ void sample_staged(const FwdStage &st, int nt, const float (&v)[VZ],
                                              float dx, const float (&zm1)[VZ], int c, int zb,
                                              float *seis, int tsamp)
{
    const int m0 = st.cp[c], m1 = st.cp[c + 1];
    for (int m = m0; m < m1; ++m) {
        const int rr = st.iz[m] - zb;
#pragma unroll
        for (int r = 0; r < VZ; ++r)
            if (rr == r)
                seis[(long)st.id[m] * nt + tsamp] = v[r] * (ABC ? sponge(dx, zm1[r]) : 1.f);
    }
}

// The three samples (steps 1-3 of a launch) of the receivers of window
// column c in this lane, written when the last step leaves the column:
// seis[rec][tsamp + k] = d * v_k for the steps k whose flag is set.
template <int VZ, bool ABC>
__device__ __forceinline__ void sample3_staged(const FwdStage &st, int nt, const float (&v1)[VZ],
                                               const float (&v2)[VZ], const float (&v3)[VZ],
                                               float dx, const float (&zm1)[VZ], int c, int zb,
                                               float *seis, int tsamp, bool r1, bool r2, bool r3)
{
    const int m0 = st.cp[c], m1 = st.cp[c + 1];
    for (int m = m0; m < m1; ++m) {
        const int rr = st.iz[m] - zb;
#pragma unroll
        for (int r = 0; r < VZ; ++r)
            if (rr == r) {
                const float d = ABC ? sponge(dx, zm1[r]) : 1.f;
                const long base = (long)st.id[m] * nt + tsamp;
                if (r1) seis[base] = v1[r] * d;
                if (r2) seis[base + 1] = v2[r] * d;
                if (r3) seis[base + 2] = v3[r] * d;
            }
    }
}

// Adds idh2 * data to the receivers of window column c in this lane, in
// table order.  Branch-free: every entry takes one FMA per row, with data 0
// for the rows that are not its own (an exact + 0).
template <int VZ, int STEP, class Stage>
__device__ __forceinline__ void inject_staged(const Stage &st, float idh2, float (&v)[VZ], int c,
                                              int zb)
{
    const int m0 = st.cp[c], m1 = st.cp[c + 1];
    for (int m = m0; m < m1; ++m) {
        const int rr = st.iz[m] - zb;
        const float val = st.v[STEP][m];
#pragma unroll
        for (int r = 0; r < VZ; ++r) v[r] = fmaf(idh2, rr == r ? val : 0.f, v[r]);
    }
}

// Decodes the wave's tile after the XCD renumbering.
struct TileId {
    int zseg, xc, slot;
};
__device__ __forceinline__ bool tile_of(int nzseg, int nxchunk, int nslots, TileId &t)
{
    const int tile = xcd_tile(nzseg * nxchunk * nslots);
    if (tile < 0) return false;
    t.zseg = tile % nzseg;
    t.xc = (tile / nzseg) % nxchunk;
    t.slot = tile / (nzseg * nxchunk);
    return true;
}

// ---------------------------------------------------------------- forward
//
// The rim tests of the forward kernels combine their terms with `&`, not
// `&&`: the short-circuit form gives the compiler a branch, and it then sinks
// each row's stencil into an exec-masked block of its own (a save/restore of
// the exec mask per row and step), which costs scalar issue slots and keeps
// the rows from being interleaved.

// One time step: u_{t+1} = leapfrog(u_t, u_{t-1}) written over u_{t-1}
// (the odd remainder of a pass, and a tuning alternative).
template <int VZ, bool ABC, bool ST, bool REC = true>
__global__ __launch_bounds__(64) void fwd1_kernel(FwdStep a, int nzseg, int nxchunk, int xw,
                                                  int nslots, int t)
{
    TileId id;
    if (!tile_of(nzseg, nxchunk, nslots, id)) return;
    const Grid g = a.g;
    const int lane = threadIdx.x;
    const int zb = 4 + id.zseg * 64 * VZ + lane * VZ;   // rows 0-3 are rim rows
    const bool active = zb < g.ldz;
    const int j0 = 4 + id.xc * xw;
    const int j1 = min(j0 + xw, g.nxx - 4);
    const int ldz = g.ldz, nxx = g.nxx;

    const long soff = (long)id.slot * g.fstride;
    const float *cur = a.cur + soff;
    const float *old = a.old + soff;
    float *out = a.out1 + soff;
    const float *cc = a.c + (long)id.slot * a.cstride;
    const int gshot = *a.shot0 + id.slot;
    const int siz = a.src_iz[gshot], six = a.src_ix[gshot];
    const float wsrc = (t + 1 < a.nt) ? a.wav[t + 1] : 0.f;
    const bool rec_on = REC && (a.seis != nullptr) && (t >= a.tlag) && (t <= a.T - 2);
    float *seis = rec_on ? a.seis + (long)id.slot * a.rec.nrec * a.nt : nullptr;
    __shared__ FwdStage st;
    constexpr bool staged = ST;     // tables fit the LDS (the host checks)
    if constexpr (ST) (void)stage_fwd(a, rec_on, ABC, j0, j1 - j0, st, lane);

    float zm1[VZ];                                   // dz - 1 of the lane's rows
    if constexpr (ABC) row_factors<VZ>(a.b1d, zb, g, zm1);

    float q[9][VZ];
#pragma unroll
    for (int k = 0; k < 9; ++k) ldcol<VZ>(q[k], cur, j0 - 4 + k, nxx, ldz, zb);
    float on[VZ], cn[VZ], pn[VZ];
    ldcol<VZ>(on, old, j0, nxx, ldz, zb);
    ldcol<VZ>(cn, cc, j0, nxx, ldz, zb);
    vzero<VZ>(pn);
    const int z0 = zb - lane * VZ;                   // the wave's first row
    float4 h = ldhalo<VZ>(cur, j0, 4, nxx - 4, ldz, z0, lane);

    for (int j = j0; j < j1; ++j) {
        float dxj = 1.f;
        if constexpr (ABC) {
            bool in;
            dxj = staged ? st.dx[j - j0] : pad_factor(a.b1d, j, nxx, g.nb, in);
        }
        float o[VZ], c[VZ];
#pragma unroll
        for (int r = 0; r < VZ; ++r) { o[r] = on[r]; c[r] = cn[r]; }
        if (j + 1 < j1) {
            ldcol<VZ>(pn, cur, j + 5, nxx, ldz, zb);
            ldcol<VZ>(on, old, j + 1, nxx, ldz, zb);
            ldcol<VZ>(cn, cc, j + 1, nxx, ldz, zb);
        }
        float ext[VZ + 8];
        zext<VZ>(ext, q[4], h);
        h = ldhalo<VZ>(cur, j + 1, 4, nxx - 4, ldz, z0, lane);
        float nv[VZ];
#pragma unroll
        for (int r = 0; r < VZ; ++r) {
            const int i = zb + r;
            const float d = ABC ? sponge(dxj, zm1[r]) : 1.f;
            float val = leapfrog<ABC>(q[4][r], o[r], c[r], lap8r<VZ, 0>(q, ext, r), d);
            nv[r] = ((i >= 4) & (i < g.nzz - 4)) ? val : o[r];
        }
        add_source<VZ>(nv, j, six, zb, siz, wsrc);
        if (active) {
            vstore<VZ>(out + (long)j * ldz + zb, nv);
            if (rec_on) {
                if constexpr (staged) sample_staged<VZ, ABC>(st, a.nt, nv, dxj, zm1, j - j0, zb, seis, t + 1 - a.tlag);
                else sample_receivers<VZ, ABC>(a, nv, j, zb, seis, t + 1 - a.tlag);
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int r = 0; r < VZ; ++r) q[k][r] = q[k + 1][r];
#pragma unroll
        for (int r = 0; r < VZ; ++r) q[8][r] = pn[r];
    }
}

// Two time steps per launch (temporal blocking).  The wave computes u_{t+1}
// on 64*VZ rows and on its x-chunk plus 4 columns each side (kept in a second
// register queue), then u_{t+2} on the inner 64*VZ - 8 rows of the chunk.
// It reads u_t, u_{t-1} and c once for both steps and writes u_{t+1} (out1)
// and u_{t+2} (out2) for the cells it owns; the four buffers are distinct,
// so no wave reads a value another wave of the same launch writes.
template <int VZ, bool ABC, bool ST, bool REC = true>
__global__ __launch_bounds__(64) void fwd2_kernel(FwdStep a, int nzseg, int nxchunk, int xw,
                                                  int nslots, int t)
{
    TileId id;
    if (!tile_of(nzseg, nxchunk, nslots, id)) return;
    const Grid g = a.g;
    const int lane = threadIdx.x;
    const int ldz = g.ldz, nxx = g.nxx;
    const int zs = id.zseg * (64 * VZ - 8);         // first computed row; owned from zs + 4
    const int zb = zs + lane * VZ;
    const bool active = zb < ldz;
    // a lane owns all or none of its rows (the bounds are multiples of VZ)
    const bool own = active && (zb >= zs + 4) && (zb < zs + 64 * VZ - 4);
    const int x0 = 4 + id.xc * xw;
    const int x1 = min(x0 + xw, nxx - 4);

    const long soff = (long)id.slot * g.fstride;
    const float *cur = a.cur + soff;
    const float *old = a.old + soff;
    float *out1 = a.out1 + soff;
    float *out2 = a.out2 + soff;
    const float *cc = a.c + (long)id.slot * a.cstride;
    const int gshot = *a.shot0 + id.slot;
    const int siz = a.src_iz[gshot], six = a.src_ix[gshot];
    const float w1 = (t + 1 < a.nt) ? a.wav[t + 1] : 0.f;
    const float w2 = (t + 2 < a.nt) ? a.wav[t + 2] : 0.f;
    const bool rec1 = REC && (a.seis != nullptr) && (t >= a.tlag) && (t <= a.T - 2);
    const bool rec2 = REC && (a.seis != nullptr) && (t + 1 >= a.tlag) && (t + 1 <= a.T - 2);
    float *seis = (rec1 || rec2) ? a.seis + (long)id.slot * a.rec.nrec * a.nt : nullptr;
    const int xa = x0 - 4;                           // window [x0 - 4, x1 + 4)
    __shared__ FwdStage st;
    constexpr bool staged = ST;     // tables fit the LDS (the host checks)
    if constexpr (ST) (void)stage_fwd(a, rec1 || rec2, ABC, xa, x1 + 4 - xa, st, lane);

    float zm1[VZ];                                   // dz - 1 of the lane's rows
    if constexpr (ABC) row_factors<VZ>(a.b1d, zb, g, zm1);

    float q1[9][VZ];     // u_t, columns j-4 .. j+4
    float q2[9][VZ];     // u_{t+1}, columns j-8 .. j
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        ldcol<VZ>(q1[k], cur, x0 - 8 + k, nxx, ldz, zb);
        vzero<VZ>(q2[k]);
    }
    float on[VZ], cn[VZ], pn[VZ];
    ldcol<VZ>(on, old, x0 - 4, nxx, ldz, zb);
    ldcol<VZ>(cn, cc, x0 - 4, nxx, ldz, zb);
    vzero<VZ>(pn);
    float4 h = ldhalo<VZ>(cur, x0 - 4, 4, nxx - 4, ldz, zs, lane);
    // c of the last five columns, each lane its own rows: step 2 at column
    // j - 4 reads what step 1 parked at column j - 4 instead of HBM
    __shared__ float cring[5][64 * VZ];
    int slot = 0;

    for (int j = x0 - 4; j < x1 + 4; ++j) {
        float o1[VZ], cj[VZ];
#pragma unroll
        for (int r = 0; r < VZ; ++r) { o1[r] = on[r]; cj[r] = cn[r]; }
        vstore<VZ>(&cring[slot][lane * VZ], cj);
        slot = (slot == 4) ? 0 : slot + 1;          // = the slot of column j - 4
        if (j + 1 < x1 + 4) {
            ldcol<VZ>(pn, cur, j + 5, nxx, ldz, zb);
            ldcol<VZ>(on, old, j + 1, nxx, ldz, zb);
            ldcol<VZ>(cn, cc, j + 1, nxx, ldz, zb);
        }
        // ---- step 1: u_{t+1} at column j (0 on the rim columns)
        float n1[VZ];
        vzero<VZ>(n1);
        if (j >= 4 && j < nxx - 4) {
            float dxj = 1.f;
            if constexpr (ABC) {
                bool in;
                dxj = staged ? st.dx[j - xa] : pad_factor(a.b1d, j, nxx, g.nb, in);
            }
            float ext[VZ + 8];
            zext<VZ>(ext, q1[4], h);
            h = ldhalo<VZ>(cur, j + 1, 4, nxx - 4, ldz, zs, lane);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const float d = ABC ? sponge(dxj, zm1[r]) : 1.f;
                float val = leapfrog<ABC>(q1[4][r], o1[r], cj[r], lap8r<VZ, 0>(q1, ext, r), d);
                n1[r] = ((i >= 4) & (i < g.nzz - 4)) ? val : o1[r];
            }
            add_source<VZ>(n1, j, six, zb, siz, w1);
            if (own && j >= x0 && j < x1) {
                vstore<VZ>(out1 + (long)j * ldz + zb, n1);
                if (rec1) {
                    if constexpr (staged) sample_staged<VZ, ABC>(st, a.nt, n1, dxj, zm1, j - xa, zb, seis, t + 1 - a.tlag);
                    else sample_receivers<VZ, ABC>(a, n1, j, zb, seis, t + 1 - a.tlag);
                }
            }
        } else {
            h = ldhalo<VZ>(cur, j + 1, 4, nxx - 4, ldz, zs, lane);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int r = 0; r < VZ; ++r) q2[k][r] = q2[k + 1][r];
#pragma unroll
        for (int r = 0; r < VZ; ++r) q2[8][r] = n1[r];

        // ---- step 2: u_{t+2} at column jj = j - 4 (an owned column)
        const int jj = j - 4;
        if (jj >= x0) {
            float dxj = 1.f;
            if constexpr (ABC) {
                bool in;
                dxj = staged ? st.dx[jj - xa] : pad_factor(a.b1d, jj, nxx, g.nb, in);
            }
            float c2[VZ];                               // c at column jj (parked 4 columns ago)
            vload<VZ>(c2, &cring[slot][lane * VZ]);
            float ext[VZ + 8];
            zext_local<VZ>(ext, q2[4]);
            float n2[VZ];
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const float o2 = q1[0][r];              // u_t at column jj
                const float d = ABC ? sponge(dxj, zm1[r]) : 1.f;
                float val = leapfrog<ABC>(q2[4][r], o2, c2[r], lap8r<VZ, 0>(q2, ext, r), d);
                n2[r] = ((i >= 4) & (i < g.nzz - 4)) ? val : o2;
            }
            add_source<VZ>(n2, jj, six, zb, siz, w2);
            if (own) {
                vstore<VZ>(out2 + (long)jj * ldz + zb, n2);
                if (rec2) {
                    if constexpr (staged) sample_staged<VZ, ABC>(st, a.nt, n2, dxj, zm1, jj - xa, zb, seis, t + 2 - a.tlag);
                    else sample_receivers<VZ, ABC>(a, n2, jj, zb, seis, t + 2 - a.tlag);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int r = 0; r < VZ; ++r) q1[k][r] = q1[k + 1][r];
#pragma unroll
        for (int r = 0; r < VZ; ++r) q1[8][r] = pn[r];
    }
}

// Three time steps per launch.  The wave computes u_{t+1} on 64*VZ rows and
// its x-chunk plus 8 columns each side, u_{t+2} on the inner 64*VZ - 8 rows
// and the chunk plus 4 columns, u_{t+3} on the inner 64*VZ - 16 rows of the
// chunk; the three fields of a column travel through three register queues.
// It reads u_t, u_{t-1} and c once and writes only what the next launch
// needs: u_{t+2} (out1) and u_{t+3} (out2) for the cells it owns.
template <int VZ, bool ABC, bool ST, bool REC = true>
__global__ __launch_bounds__(64, 3) void fwd3_kernel(FwdStep a, int nzseg, int nxchunk, int xw,
                                                  int nslots, int t)
{
    TileId id;
    if (!tile_of(nzseg, nxchunk, nslots, id)) return;
    const Grid g = a.g;
    const int lane = threadIdx.x;
    const int ldz = g.ldz, nxx = g.nxx;
    const int zs = id.zseg * (64 * VZ - 16) - 4;    // first computed row; owned from zs + 8
    const int zb = zs + lane * VZ;
    const int zl = zb < 0 ? 0 : zb;                  // load row: rows above 0 feed rim rows only
    const unsigned uzl = unsigned(zl);
    const bool active = zb < ldz;
    const bool own = active && (zb >= zs + 8) && (zb < zs + 64 * VZ - 8);
    const int x0 = 4 + id.xc * xw;
    const int x1 = min(x0 + xw, nxx - 4);
    // odd chunks march right to left: neighbouring chunks then read their
    // shared halo columns at the same time (column c of the march is PC(c))
    const bool back = (id.xc & 1) != 0;
    const int mir = x0 + x1 - 1;
    auto PC = [&](int c) { return back ? mir - c : c; };

    const long soff = (long)id.slot * g.fstride;
    const float *cur = a.cur + soff;
    const float *old = a.old + soff;
    float *out1 = a.out1 + soff;
    float *out2 = a.out2 + soff;
    const float *cc = a.c + (long)id.slot * a.cstride;
    const int gshot = *a.shot0 + id.slot;
    const int siz = a.src_iz[gshot], six = a.src_ix[gshot];
    const float w1 = (t + 1 < a.nt) ? a.wav[t + 1] : 0.f;
    const float w2 = (t + 2 < a.nt) ? a.wav[t + 2] : 0.f;
    const float w3 = (t + 3 < a.nt) ? a.wav[t + 3] : 0.f;
    const bool rec1 = REC && (a.seis != nullptr) && (t >= a.tlag) && (t <= a.T - 2);
    const bool rec2 = REC && (a.seis != nullptr) && (t + 1 >= a.tlag) && (t + 1 <= a.T - 2);
    const bool rec3 = REC && (a.seis != nullptr) && (t + 2 >= a.tlag) && (t + 2 <= a.T - 2);
    float *seis = (rec1 || rec2 || rec3) ? a.seis + (long)id.slot * a.rec.nrec * a.nt : nullptr;
    const int xa = x0 - 8;                           // window [x0 - 8, x1 + 8)
    __shared__ FwdStage st;
    constexpr bool staged = ST;     // tables fit the LDS (the host checks)
    if constexpr (ST) (void)stage_fwd(a, rec1 || rec2 || rec3, ABC, xa, x1 + 8 - xa, st, lane);

    float zm1[VZ];                                   // dz - 1 of the lane's rows
    if constexpr (ABC) row_factors<VZ>(a.b1d, zb, g, zm1);

    float q1[9][VZ];     // u_t
    float q2[9][VZ];     // u_{t+1}
    float q3[9][VZ];     // u_{t+2}
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        ldcolc<VZ>(q1[m], cur, PC(x0 - 12 + m), nxx, ldz, zl);
        vzero<VZ>(q2[m]);
        vzero<VZ>(q3[m]);
    }
    // halos (ldhalo1) of the u_t queue columns, in the queue's slots
    float hq[9];
#pragma unroll
    for (int m = 4; m < 9; ++m) hq[m] = ldhalo1<VZ>(cur, PC(x0 - 12 + m), nxx, ldz, zs, lane);
    hq[0] = hq[1] = hq[2] = hq[3] = 0.f;
    float on[VZ], cn[VZ], pn[VZ], hn;
    ldcolc<VZ>(on, old, PC(x0 - 8), nxx, ldz, zl);
    ldcolc<VZ>(cn, cc, PC(x0 - 8), nxx, ldz, zl);
    __shared__ float cring[9][64 * VZ];              // c of the last nine columns
    const int jend = x1 + 8;
    int j = x0 - 8;

    // One column j; P is the phase of the unrolled loop: slot (m + P) % 9 of
    // a queue holds its column m (q1: j-4+m; q2, cring: j-8+m; q3: j-12+m
    // once the column's results are in).  Steps 1 and 2 run on every column:
    // a column outside [4, nxx-4) takes its old value, a zero rim value.
    auto column = [&](auto phase) {
        constexpr int P = decltype(phase)::value;
        constexpr int C0 = P % 9, C4 = (P + 4) % 9, C8 = (P + 8) % 9;
        float o1[VZ], cj[VZ];
#pragma unroll
        for (int r = 0; r < VZ; ++r) { o1[r] = on[r]; cj[r] = cn[r]; }
        vstore<VZ>(&cring[C8][lane * VZ], cj);
        ldcolc<VZ>(pn, cur, PC(j + 5), nxx, ldz, zl);
        hn = ldhalo1<VZ>(cur, PC(j + 5), nxx, ldz, zs, lane);   // beside its main load
        ldcolc<VZ>(on, old, PC(j + 1), nxx, ldz, zl);
        ldcolc<VZ>(cn, cc, PC(j + 1), nxx, ldz, zl);
        // ---- step 1: u_{t+1} at column j
        {
            const int pj = PC(j);
            const bool jin = (pj >= 4) && (pj < nxx - 4);
            float dxj = 1.f;
            if constexpr (ABC) {
                bool in;
                dxj = staged ? st.dx[pj - xa] : pad_factor(a.b1d, pj, nxx, g.nb, in);
            }
            float ext[VZ + 8], dz[VZ + 1];
            zext1<VZ>(ext, q1[C4], hq[C4]);
            zdiff<VZ>(dz, ext);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const float d = ABC ? sponge(dxj, zm1[r]) : 1.f;
                float val = leapfrog<ABC>(q1[C4][r], o1[r], cj[r], lap8z<VZ, C0>(q1, ext, dz, r), d);
                q2[C8][r] = (jin & (i >= 4) & (i < g.nzz - 4)) ? val : o1[r];
            }
            add_source<VZ>(q2[C8], pj, six, zb, siz, w1);
        }
        // ---- step 2: u_{t+2} at column jj = j - 4, from x0 - 4 on (step 3
        // reads no column before it; the queue slots of the skipped columns
        // are unused)
        if (j - 4 >= x0 - 4) {
            const int jj = j - 4, pjj = PC(jj);
            const bool jin = (pjj >= 4) && (pjj < nxx - 4);
            float dxj = 1.f;
            if constexpr (ABC) {
                bool in;
                // columns outside the window (the march's first four) are not used
                dxj = staged ? st.dx[min(max(pjj - xa, 0), x1 + 7 - xa)]
                             : pad_factor(a.b1d, pjj, nxx, g.nb, in);
            }
            float ext[VZ + 8], c2v[VZ], dz[VZ + 1];
            zext_local<VZ>(ext, q2[C4]);
            vload<VZ>(c2v, &cring[C4][lane * VZ]);
            zdiff<VZ>(dz, ext);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const float o2 = q1[C0][r];             // u_t at column jj
                const float c2 = c2v[r];
                const float d = ABC ? sponge(dxj, zm1[r]) : 1.f;
                float val = leapfrog<ABC>(q2[C4][r], o2, c2, lap8z<VZ, C0>(q2, ext, dz, r), d);
                q3[C8][r] = (jin & (i >= 4) & (i < g.nzz - 4)) ? val : o2;
            }
            add_source<VZ>(q3[C8], pjj, six, zb, siz, w2);
            if (own && jj >= x0 && jj < x1) vstore<VZ>((out1 + (long)pjj * ldz) + uzl, q3[C8]);
        }
        // ---- step 3: u_{t+3} at column jjj = j - 8 (an owned column)
        const int jjj = j - 8;
        if (jjj >= x0) {
            const int pjjj = PC(jjj);
            float dxj = 1.f;
            if constexpr (ABC) {
                bool in;
                dxj = staged ? st.dx[pjjj - xa] : pad_factor(a.b1d, pjjj, nxx, g.nb, in);
            }
            float c3[VZ];
            vload<VZ>(c3, &cring[C0][lane * VZ]);
            float ext[VZ + 8], dz[VZ + 1];
            zext_local<VZ>(ext, q3[C4]);
            zdiff<VZ>(dz, ext);
            float n3[VZ];
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const float o3 = q2[C0][r];             // u_{t+1} at column jjj
                const float d = ABC ? sponge(dxj, zm1[r]) : 1.f;
                float val = leapfrog<ABC>(q3[C4][r], o3, c3[r], lap8z<VZ, C0>(q3, ext, dz, r), d);
                n3[r] = ((i >= 4) & (i < g.nzz - 4)) ? val : o3;
            }
            add_source<VZ>(n3, pjjj, six, zb, siz, w3);
            if (own) {
                vstore<VZ>((out2 + (long)pjjj * ldz) + uzl, n3);
                // the column's three samples: u_{t+1} (q2), u_{t+2} (q3), u_{t+3}
                if (rec1 || rec2 || rec3) {
                    if constexpr (staged)
                        sample3_staged<VZ, ABC>(st, a.nt, q2[C0], q3[C4], n3, dxj, zm1, pjjj - xa, zb,
                                                seis, t + 1 - a.tlag, rec1, rec2, rec3);
                    else
                        sample3_receivers<VZ, ABC>(a, q2[C0], q3[C4], n3, pjjj, zb, seis,
                                                   t + 1 - a.tlag, rec1, rec2, rec3);
                }
            }
        }
        // column j + 5 of u_t replaces column j - 4
#pragma unroll
        for (int r = 0; r < VZ; ++r) q1[C0][r] = pn[r];
        hq[C0] = hn;
    };
    for (;;) {
        column(Phase<0>{}); if (++j >= jend) break;
        column(Phase<1>{}); if (++j >= jend) break;
        column(Phase<2>{}); if (++j >= jend) break;
        column(Phase<3>{}); if (++j >= jend) break;
        column(Phase<4>{}); if (++j >= jend) break;
        column(Phase<5>{}); if (++j >= jend) break;
        column(Phase<6>{}); if (++j >= jend) break;
        column(Phase<7>{}); if (++j >= jend) break;
        column(Phase<8>{}); if (++j >= jend) break;
    }
}

// Two samples (steps s and s+1 of a launch) of the receivers of window
// column c in this lane: seis[rec][tsamp + k] = d * v_k for the flagged k.
template <int VZ, bool ABC>
__device__ __forceinline__ void sample2_staged(const FwdStage &st, int nt, const float (&v1)[VZ],
                                               const float (&v2)[VZ], float dx,
                                               const float (&zm1)[VZ], int c, int zb,
                                               float *seis, int tsamp, bool r1, bool r2)
{
    const int m0 = st.cp[c], m1 = st.cp[c + 1];
    for (int m = m0; m < m1; ++m) {
        const int rr = st.iz[m] - zb;
#pragma unroll
        for (int r = 0; r < VZ; ++r)
            if (rr == r) {
                const float d = ABC ? sponge(dx, zm1[r]) : 1.f;
                const long base = (long)st.id[m] * nt + tsamp;
                if (r1) seis[base] = v1[r] * d;
                if (r2) seis[base + 1] = v2[r] * d;
            }
    }
}

// sample2_staged from the global receiver table.
template <int VZ, bool ABC>
__device__ __forceinline__ void sample2_receivers(const FwdStep &a, const float (&v1)[VZ],
                                                  const float (&v2)[VZ], int j, int zb,
                                                  float *seis, int tsamp, bool r1, bool r2)
{
    const int k0 = a.rec.col_ptr[j], k1 = a.rec.col_ptr[j + 1];
    for (int k = k0; k < k1; ++k) {
        const int rr = a.rec.iz[k] - zb;
#pragma unroll
        for (int r = 0; r < VZ; ++r)
            if (rr == r) {
                const float d = ABC ? damper(a.b1d, zb + r, j, a.g) : 1.f;
                const long base = (long)a.rec.id[k] * a.nt + tsamp;
                if (r1) seis[base] = v1[r] * d;
                if (r2) seis[base + 1] = v2[r] * d;
            }
    }
}

// Four time steps per launch: the three-step march with a fourth register
// queue.  Step s runs on the chunk +- 4(4 - s) columns, on the rows that
// stay valid without the neighbouring waves' values (64*VZ rows for step 1,
// 8 fewer for each later step; 64*VZ - 24 owned), so a launch reads u_t,
// u_{t-1} and c once and writes u_{t+3} and u_{t+4}: the HBM bytes of four
// steps are those of three in fwd3_kernel.  c of the last thirteen columns
// waits in LDS for steps 2-4.  A receiver's samples of steps 1 and 2 are
// written when step 2 reaches its column (u_{t+1} is then still in the
// second queue), those of steps 3 and 4 when step 4 does.  Same operations
// in the same order as every other forward kernel, so the same bits.
template <int VZ, bool ABC, bool ST, bool REC = true>
__global__ __launch_bounds__(64) void fwd4_kernel(FwdStep a, int nzseg, int nxchunk, int xw,
                                                  int nslots, int t)
{
    TileId id;
    if (!tile_of(nzseg, nxchunk, nslots, id)) return;
    const Grid g = a.g;
    const int lane = threadIdx.x;
    const int ldz = g.ldz, nxx = g.nxx;
    const int zs = id.zseg * (64 * VZ - 24) - 8;    // first computed row; owned from zs + 12
    const int zb = zs + lane * VZ;
    const int zl = zb < 0 ? 0 : zb;
    const unsigned uzl = unsigned(zl);
    const bool active = zb < ldz;
    const bool own = active && (zb >= zs + 12) && (zb < zs + 64 * VZ - 12);
    const int x0 = 4 + id.xc * xw;
    const int x1 = min(x0 + xw, nxx - 4);
    const bool back = (id.xc & 1) != 0;
    const int mir = x0 + x1 - 1;
    auto PC = [&](int c) { return back ? mir - c : c; };

    const long soff = (long)id.slot * g.fstride;
    const float *cur = a.cur + soff;
    const float *old = a.old + soff;
    float *out1 = a.out1 + soff;
    float *out2 = a.out2 + soff;
    const float *cc = a.c + (long)id.slot * a.cstride;
    const int gshot = *a.shot0 + id.slot;
    const int siz = a.src_iz[gshot], six = a.src_ix[gshot];
    const float w1 = (t + 1 < a.nt) ? a.wav[t + 1] : 0.f;
    const float w2 = (t + 2 < a.nt) ? a.wav[t + 2] : 0.f;
    const float w3 = (t + 3 < a.nt) ? a.wav[t + 3] : 0.f;
    const float w4 = (t + 4 < a.nt) ? a.wav[t + 4] : 0.f;
    const bool rec1 = REC && (a.seis != nullptr) && (t >= a.tlag) && (t <= a.T - 2);
    const bool rec2 = REC && (a.seis != nullptr) && (t + 1 >= a.tlag) && (t + 1 <= a.T - 2);
    const bool rec3 = REC && (a.seis != nullptr) && (t + 2 >= a.tlag) && (t + 2 <= a.T - 2);
    const bool rec4 = REC && (a.seis != nullptr) && (t + 3 >= a.tlag) && (t + 3 <= a.T - 2);
    float *seis = (rec1 || rec2 || rec3 || rec4) ? a.seis + (long)id.slot * a.rec.nrec * a.nt
                                                 : nullptr;
    const int xa = x0 - 12;                          // window [x0 - 12, x1 + 12)
    __shared__ FwdStage st;
    constexpr bool staged = ST;     // tables fit the LDS (the host checks)
    if constexpr (ST)
        (void)stage_fwd(a, rec1 || rec2 || rec3 || rec4, ABC, xa, x1 + 12 - xa, st, lane);

    float zm1[VZ];                                   // dz - 1 of the lane's rows
    if constexpr (ABC) row_factors<VZ>(a.b1d, zb, g, zm1);
    // sponge x factor of grid column pc (1 without the sponge)
    auto xfac = [&](int pc) {
        if constexpr (ABC) {
            bool in;
            return staged ? st.dx[min(max(pc - xa, 0), x1 + 11 - xa)]
                          : pad_factor(a.b1d, pc, nxx, g.nb, in);
        } else {
            return 1.f;
        }
    };

    float q1[9][VZ];     // u_t
    float q2[9][VZ];     // u_{t+1}
    float q3[9][VZ];     // u_{t+2}
    float q4[9][VZ];     // u_{t+3}
    float hq[9];         // halos (ldhalo1) of the q1 columns
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        ldcolc<VZ>(q1[m], cur, PC(x0 - 16 + m), nxx, ldz, zl);
        vzero<VZ>(q2[m]);
        vzero<VZ>(q3[m]);
        vzero<VZ>(q4[m]);
    }
#pragma unroll
    for (int m = 4; m < 9; ++m) hq[m] = ldhalo1<VZ>(cur, PC(x0 - 16 + m), nxx, ldz, zs, lane);
    hq[0] = hq[1] = hq[2] = hq[3] = 0.f;
    float on[VZ], cn[VZ], pn[VZ], hn;
    ldcolc<VZ>(on, old, PC(x0 - 12), nxx, ldz, zl);
    ldcolc<VZ>(cn, cc, PC(x0 - 12), nxx, ldz, zl);
    __shared__ float cring[13][64 * VZ];             // c of the last thirteen columns
    const int jend = x1 + 12;
    int j = x0 - 12;
    int cs = 0;                                      // cring slot of column j

    // One column j; slot (m + P) % 9 of a queue holds its column m (q1:
    // j-4+m; q2: j-8+m; q3: j-12+m; q4: j-16+m once the column's results
    // are in).  Steps 1-3 keep the old (zero rim) value outside [4, nxx-4).
    auto column = [&](auto phase) {
        constexpr int P = decltype(phase)::value;
        constexpr int C0 = P % 9, C4 = (P + 4) % 9, C8 = (P + 8) % 9;
        float o1[VZ], cj[VZ];
#pragma unroll
        for (int r = 0; r < VZ; ++r) { o1[r] = on[r]; cj[r] = cn[r]; }
        vstore<VZ>(&cring[cs][lane * VZ], cj);
        ldcolc<VZ>(pn, cur, PC(j + 5), nxx, ldz, zl);
        hn = ldhalo1<VZ>(cur, PC(j + 5), nxx, ldz, zs, lane);
        ldcolc<VZ>(on, old, PC(j + 1), nxx, ldz, zl);
        ldcolc<VZ>(cn, cc, PC(j + 1), nxx, ldz, zl);
        const int cs2 = cs >= 4 ? cs - 4 : cs + 9;       // slot of column j - 4
        const int cs3 = cs >= 8 ? cs - 8 : cs + 5;       // j - 8
        const int cs4 = cs == 12 ? 0 : cs + 1;           // j - 12
        // ---- step 1: u_{t+1} at column j
        {
            const int pj = PC(j);
            const bool jin = (pj >= 4) && (pj < nxx - 4);
            const float dxj = xfac(pj);
            float ext[VZ + 8], dz[VZ + 1];
            zext1<VZ>(ext, q1[C4], hq[C4]);
            zdiff<VZ>(dz, ext);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const float d = ABC ? sponge(dxj, zm1[r]) : 1.f;
                float val = leapfrog<ABC>(q1[C4][r], o1[r], cj[r], lap8z<VZ, C0>(q1, ext, dz, r), d);
                q2[C8][r] = (jin & (i >= 4) & (i < g.nzz - 4)) ? val : o1[r];
            }
            add_source<VZ>(q2[C8], pj, six, zb, siz, w1);
        }
        // ---- step 2: u_{t+2} at column j - 4, from x0 - 8 on
        if (j - 4 >= x0 - 8) {
            const int jj = j - 4, pjj = PC(jj);
            const bool jin = (pjj >= 4) && (pjj < nxx - 4);
            const float dxj = xfac(pjj);
            float ext[VZ + 8], c2v[VZ], dz[VZ + 1];
            zext_local<VZ>(ext, q2[C4]);
            vload<VZ>(c2v, &cring[cs2][lane * VZ]);
            zdiff<VZ>(dz, ext);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const float o2 = q1[C0][r];             // u_t at column jj
                const float d = ABC ? sponge(dxj, zm1[r]) : 1.f;
                float val = leapfrog<ABC>(q2[C4][r], o2, c2v[r], lap8z<VZ, C0>(q2, ext, dz, r), d);
                q3[C8][r] = (jin & (i >= 4) & (i < g.nzz - 4)) ? val : o2;
            }
            add_source<VZ>(q3[C8], pjj, six, zb, siz, w2);
            // samples of steps 1 and 2: u_{t+1} (q2) and u_{t+2} of an owned column
            if (own && jj >= x0 && jj < x1 && (rec1 || rec2)) {
                if constexpr (staged)
                    sample2_staged<VZ, ABC>(st, a.nt, q2[C4], q3[C8], dxj, zm1, pjj - xa, zb,
                                            seis, t + 1 - a.tlag, rec1, rec2);
                else
                    sample2_receivers<VZ, ABC>(a, q2[C4], q3[C8], pjj, zb, seis,
                                               t + 1 - a.tlag, rec1, rec2);
            }
        }
        // ---- step 3: u_{t+3} at column j - 8, from x0 - 4 on
        if (j - 8 >= x0 - 4) {
            const int jjj = j - 8, pjjj = PC(jjj);
            const bool jin = (pjjj >= 4) && (pjjj < nxx - 4);
            const float dxj = xfac(pjjj);
            float ext[VZ + 8], c3v[VZ], dz[VZ + 1];
            zext_local<VZ>(ext, q3[C4]);
            vload<VZ>(c3v, &cring[cs3][lane * VZ]);
            zdiff<VZ>(dz, ext);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const float o3 = q2[C0][r];             // u_{t+1} at column jjj
                const float d = ABC ? sponge(dxj, zm1[r]) : 1.f;
                float val = leapfrog<ABC>(q3[C4][r], o3, c3v[r], lap8z<VZ, C0>(q3, ext, dz, r), d);
                q4[C8][r] = (jin & (i >= 4) & (i < g.nzz - 4)) ? val : o3;
            }
            add_source<VZ>(q4[C8], pjjj, six, zb, siz, w3);
            if (own && jjj >= x0 && jjj < x1) vstore<VZ>((out1 + (long)pjjj * ldz) + uzl, q4[C8]);
        }
        // ---- step 4: u_{t+4} at column j - 12 (an owned column)
        const int j4 = j - 12;
        if (j4 >= x0) {
            const int pj4 = PC(j4);
            const float dxj = xfac(pj4);
            float c4v[VZ], ext[VZ + 8], dz[VZ + 1], n4[VZ];
            vload<VZ>(c4v, &cring[cs4][lane * VZ]);
            zext_local<VZ>(ext, q4[C4]);
            zdiff<VZ>(dz, ext);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const float o4 = q3[C0][r];             // u_{t+2} at column j4
                const float d = ABC ? sponge(dxj, zm1[r]) : 1.f;
                float val = leapfrog<ABC>(q4[C4][r], o4, c4v[r], lap8z<VZ, C0>(q4, ext, dz, r), d);
                n4[r] = ((i >= 4) & (i < g.nzz - 4)) ? val : o4;
            }
            add_source<VZ>(n4, pj4, six, zb, siz, w4);
            if (own) {
                vstore<VZ>((out2 + (long)pj4 * ldz) + uzl, n4);
                // samples of steps 3 and 4: u_{t+3} (q4) and u_{t+4}
                if (rec3 || rec4) {
                    if constexpr (staged)
                        sample2_staged<VZ, ABC>(st, a.nt, q4[C4], n4, dxj, zm1, pj4 - xa, zb,
                                                seis, t + 3 - a.tlag, rec3, rec4);
                    else
                        sample2_receivers<VZ, ABC>(a, q4[C4], n4, pj4, zb, seis,
                                                   t + 3 - a.tlag, rec3, rec4);
                }
            }
        }
        // column j + 5 of u_t replaces column j - 4
#pragma unroll
        for (int r = 0; r < VZ; ++r) q1[C0][r] = pn[r];
        hq[C0] = hn;
        cs = (cs == 12) ? 0 : cs + 1;
    };
    for (;;) {
        column(Phase<0>{}); if (++j >= jend) break;
        column(Phase<1>{}); if (++j >= jend) break;
        column(Phase<2>{}); if (++j >= jend) break;
        column(Phase<3>{}); if (++j >= jend) break;
        column(Phase<4>{}); if (++j >= jend) break;
        column(Phase<5>{}); if (++j >= jend) break;
        column(Phase<6>{}); if (++j >= jend) break;
        column(Phase<7>{}); if (++j >= jend) break;
        column(Phase<8>{}); if (++j >= jend) break;
    }
}

// ---------------------------------------------------------------- 2-D tiles
//
// Forward steps on 64 x 64-cell tiles held in LDS, one workgroup of NW
// waves per tile (each thread: one row of every NW-th column): for batches
// of one or a few shots, where the column-marching waves above are too few
// to fill the device and each wave's march over its chunk plus halo sets the
// launch time.  A launch of NS steps loads u_t,
// u_{t-1} and c of the tile (out-of-grid cells as zero) and runs the steps
// in LDS, each on a region 4 cells narrower on every side than the last; the
// inner (64 - 8 NS)^2 cells are the tile's own.  Every cell is updated with
// the arithmetic of the column kernels (lap8_cell is lap8r term for term,
// the same leapfrog, rim select, source and sample formulas), so the fields
// and seismograms are bit-identical to theirs.

// lap8r of cell (x, z) of an LDS tile, term for term: the x neighbours from
// LDS, the z neighbours shifted in from the lanes above and below (the
// thread of lane z holds row z of the column; every lane of the wave must
// run this, and only rows 4 .. 59 get all eight).
__device__ __forceinline__ float lap8_cell(const float (*u)[kTile], int x, int z, float p)
{
    const float a1 = wave_shr1z(p), a2 = wave_shr1z(a1), a3 = wave_shr1z(a2), a4 = wave_shr1z(a3);
    const float b1 = wave_shl1z(p), b2 = wave_shl1z(b1), b3 = wave_shl1z(b2), b4 = wave_shl1z(b3);
    const float s1 = ((u[x - 1][z] - p) + (u[x + 1][z] - p)) + ((a1 - p) + (b1 - p));
    const float s2 = fmaf(-4.f, p, (u[x - 2][z] + u[x + 2][z]) + (a2 + b2));
    const float s3 = fmaf(-4.f, p, (u[x - 3][z] + u[x + 3][z]) + (a3 + b3));
    const float s4 = fmaf(-4.f, p, (u[x - 4][z] + u[x + 4][z]) + (a4 + b4));
    return fmaf(kC4, s1, fmaf(-kC3, s2, fmaf(kC2, s3, -kC1 * s4)));
}

// Tiles of a launch of `steps` steps: owned (64 - 8 steps)^2 cells each over
// the updated rows and columns 4 .. n-5.
__host__ __device__ __forceinline__ int tile_count(int n, int steps)
{
    const int own = kTile - 8 * steps;
    const int k = (n - 8 + own - 1) / own;
    return k > 1 ? k : 1;
}

template <int NS, bool ABC, int NW>
__global__ __launch_bounds__(64 * NW) void fwdt_kernel(FwdStep a, int ntz, int ntx, int nslots,
                                                       int t)
{
    constexpr int H = 4 * NS, OWN = kTile - 2 * H, CPT = kTile / NW;   // columns per thread
    const int tile = xcd_tile(ntz * ntx * nslots);
    if (tile < 0) return;                            // the whole workgroup
    const int tz = tile % ntz, tx = (tile / ntz) % ntx, slot = tile / (ntz * ntx);
    const Grid g = a.g;
    const int ldz = g.ldz, nxx = g.nxx, nzz = g.nzz;
    const int z0 = 4 + tz * OWN - H;                 // grid row of tile row 0
    const int x0 = 4 + tx * OWN - H;                 // grid column of tile column 0
    const int z = threadIdx.x & 63;                  // tile row of this thread
    const int xg = threadIdx.x >> 6;                 // its columns: xg, xg + NW, ...
    const int gz = z0 + z;
    const bool zok = (gz >= 0) & (gz < nzz);
    const long soff = (long)slot * g.fstride;
    const float *cur = a.cur + soff;
    const float *old = a.old + soff;
    const float *cc = a.c + (long)slot * a.cstride;
    const int gshot = *a.shot0 + slot;
    const int siz = a.src_iz[gshot], six = a.src_ix[gshot];

    __shared__ float u[2][kTile][kTile];             // [buffer][column][row]
    __shared__ float cs[kTile][kTile];
    __shared__ float sdx[kTile], szm[kTile];         // sponge: dx of a column, dz - 1 of a row
    __shared__ int scp[kTile + 1];                   // receiver table offsets of the own columns
    const int zlo = 4 + tz * OWN, xlo = 4 + tx * OWN;    // own rows / columns in the grid
    const int zhi = min(zlo + OWN, nzz), xhi = min(xlo + OWN, nxx - 4);
    const bool rec_any = (a.seis != nullptr) && (t + NS - 1 >= a.tlag) && (t <= a.T - 2);
    if (rec_any && (int)threadIdx.x <= OWN)
        scp[threadIdx.x] = a.rec.col_ptr[min(xlo + (int)threadIdx.x, max(xhi, xlo))];
    if constexpr (ABC) {
        bool in;
        if (threadIdx.x < kTile) sdx[z] = pad_factor(a.b1d, x0 + z, nxx, g.nb, in);
        else if (threadIdx.x < 2 * kTile) szm[z] = pad_factor(a.b1d, gz, nzz, g.nb, in) - 1.f;
    }
    {
        // all loads issued before the first LDS store, so that their
        // latencies overlap
        float lu[CPT], lo[CPT], lc[CPT];
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const int x = xg + NW * m, gx = x0 + x;
            const bool ok = zok & (gx >= 0) & (gx < nxx);
            const long idx = (long)gx * ldz + gz;
            lu[m] = ok ? cur[idx] : 0.f;
            lc[m] = ok ? cc[idx] : 0.f;
            // u_{t-1} only where step 1 reads it (with one step: the own
            // cells, which no other workgroup reads while the result
            // overwrites them)
            const bool need = (x >= 4) & (x < kTile - 4) & (z >= 4) & (z < kTile - 4);
            lo[m] = (ok & need) ? old[idx] : 0.f;
        }
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const int x = xg + NW * m;
            u[0][x][z] = lu[m];
            u[1][x][z] = lo[m];
            cs[x][z] = lc[m];
        }
    }
    __syncthreads();

#pragma unroll
    for (int k = 1; k <= NS; ++k) {
        const float (*cu)[kTile] = u[(k + 1) & 1];       // u_{t+k-1}
        float (*nu)[kTile] = u[k & 1];                   // u_{t+k-2} -> u_{t+k}
        const float w = (t + k < a.nt) ? a.wav[t + k] : 0.f;
        const bool zin = (z >= 4 * k) & (z < kTile - 4 * k);
        const float dzm = ABC ? szm[z] : 0.f;
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const int x = xg + NW * m;                   // the same for the whole wave
            if ((x < 4 * k) | (x >= kTile - 4 * k)) continue;
            const float p = cu[x][z];
            const float lap = lap8_cell(cu, x, z, p);    // all 64 lanes
            if (zin) {
                const int gx = x0 + x;
                const float o = nu[x][z];
                const float d = ABC ? sponge(sdx[x], dzm) : 1.f;
                const float val = leapfrog<ABC>(p, o, cs[x][z], lap, d);
                float v = ((gx >= 4) & (gx < nxx - 4) & (gz >= 4) & (gz < nzz - 4)) ? val : o;
                if ((gx == six) & (gz == siz)) v += w;
                nu[x][z] = v;
            }
        }
        __syncthreads();
        // the samples of u_{t+k} at the receivers of the own cells
        const bool rec = (a.seis != nullptr) && (t + k - 1 >= a.tlag) && (t + k - 1 <= a.T - 2);
        if (rec && (int)threadIdx.x < xhi - xlo) {       // one own column per thread
            float *seis = a.seis + (long)slot * a.rec.nrec * a.nt;
            const int c = threadIdx.x, ix = xlo + c;
            for (int e = scp[c]; e < scp[c + 1]; ++e) {
                const int iz = a.rec.iz[e];
                if (iz < zlo || iz >= zhi) continue;
                const float v = nu[ix - x0][iz - z0];
                const float d = ABC ? sponge(sdx[ix - x0], szm[iz - z0]) : 1.f;
                seis[(long)a.rec.id[e] * a.nt + (t + k - a.tlag)] = v * d;
            }
        }
        // u_{t+NS-1} to out1 and u_{t+NS} to out2 (one step: u_{t+1} to out1)
        float *out = (k == NS) ? (NS == 1 ? a.out1 : a.out2) : (k == NS - 1 ? a.out1 : nullptr);
        if (out && (gz >= zlo) & (gz < zhi)) {
            out += soff;
#pragma unroll
            for (int m = 0; m < CPT; ++m) {
                const int x = xg + NW * m, gx = x0 + x;
                if ((gx >= xlo) & (gx < xhi)) out[(long)gx * ldz + gz] = nu[x][z];
            }
        }
    }
}

// ---------------------------------------------------------------- reverse

// One reverse step k: S_new = leapfrog(S_cur, S_old) over S_old, R likewise,
// data injected into R_{k+1}, illum += S_new^2, image += S_new R_k or
// gradient += R_k c lap(S_cur).
template <int VZ, int MODE, bool ST>
__global__ __launch_bounds__(64) void rev1_kernel(RevStep a, int nzseg, int nxchunk, int xw,
                                                  int nslots, int k)
{
    TileId id;
    if (!tile_of(nzseg, nxchunk, nslots, id)) return;
    const Grid g = a.g;
    const int lane = threadIdx.x;
    const int zb = 4 + id.zseg * 64 * VZ + lane * VZ;   // rows 0-3 are rim rows
    const bool active = zb < g.ldz;
    const int j0 = 4 + id.xc * xw;
    const int j1 = min(j0 + xw, g.nxx - 4);
    const int ldz = g.ldz, nxx = g.nxx;

    const long soff = (long)id.slot * g.fstride;
    const float *scur = a.scur + soff;
    const float *sold = a.sold + soff;
    float *sout = a.sout1 + soff;
    const float *rcur = a.rcur + soff;
    const float *rold = a.rold + soff;
    float *rout = a.rout1 + soff;
    const float *cc = a.c + (long)id.slot * a.cstride;
    float *acc = a.acc + soff;
    float *ill = a.ill + soff;
    const bool inj_on = (k + 1 < a.nt);
    const int drow = a.data_by_shot ? (*a.shot0 + id.slot) : id.slot;
    const float *dat = a.data + (long)drow * a.rec.nrec * a.nt;
    const int zw = 4 + id.zseg * 64 * VZ;
    __shared__ RevStage st;
    constexpr bool staged = ST;     // tables fit the LDS (the host checks)
    if constexpr (ST) (void)stage_rev(a, dat, j0, j1 - j0, zw, zw + 64 * VZ, inj_on, a.nt - 2 - k,
                                        false, 0, false, 0, st, lane);

    float qs[9][VZ], qr[9][VZ];
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        ldcol<VZ>(qs[m], scur, j0 - 4 + m, nxx, ldz, zb);
        ldcol<VZ>(qr[m], rcur, j0 - 4 + m, nxx, ldz, zb);
    }
    float osn[VZ], orn[VZ], cn[VZ], an[VZ], in[VZ], psn[VZ], prn[VZ];
    ldcol<VZ>(osn, sold, j0, nxx, ldz, zb);
    ldcol<VZ>(orn, rold, j0, nxx, ldz, zb);
    ldcol<VZ>(cn, cc, j0, nxx, ldz, zb);
    ldcol<VZ>(an, acc, j0, nxx, ldz, zb);
    ldcol<VZ>(in, ill, j0, nxx, ldz, zb);
    vzero<VZ>(psn); vzero<VZ>(prn);
    float4 hs = ldhalo<VZ>(scur, j0, 4, nxx - 4, ldz, zw, lane);
    float4 hr = ldhalo<VZ>(rcur, j0, 4, nxx - 4, ldz, zw, lane);

    for (int j = j0; j < j1; ++j) {
        float os[VZ], orr[VZ], c[VZ], ac[VZ], il[VZ];
#pragma unroll
        for (int r = 0; r < VZ; ++r) {
            os[r] = osn[r]; orr[r] = orn[r]; c[r] = cn[r]; ac[r] = an[r]; il[r] = in[r];
        }
        if (j + 1 < j1) {
            ldcol<VZ>(psn, scur, j + 5, nxx, ldz, zb);
            ldcol<VZ>(prn, rcur, j + 5, nxx, ldz, zb);
            ldcol<VZ>(osn, sold, j + 1, nxx, ldz, zb);
            ldcol<VZ>(orn, rold, j + 1, nxx, ldz, zb);
            ldcol<VZ>(cn, cc, j + 1, nxx, ldz, zb);
            ldcol<VZ>(an, acc, j + 1, nxx, ldz, zb);
            ldcol<VZ>(in, ill, j + 1, nxx, ldz, zb);
        }
        float es[VZ + 8], er[VZ + 8];
        zext<VZ>(es, qs[4], hs);
        zext<VZ>(er, qr[4], hr);
        hs = ldhalo<VZ>(scur, j + 1, 4, nxx - 4, ldz, zw, lane);
        hr = ldhalo<VZ>(rcur, j + 1, 4, nxx - 4, ldz, zw, lane);
        float ns[VZ], nr[VZ];
#pragma unroll
        for (int r = 0; r < VZ; ++r) {
            const int i = zb + r;
            const bool inside = (i >= 4) && (i < g.nzz - 4);
            const float pr = qr[4][r];
            const float lapS = lap8r<VZ, 0>(qs, es, r);
            const float sn = fmaf(c[r], lapS, qs[4][r] + (qs[4][r] - os[r]));
            const float rn = fmaf(c[r], lap8r<VZ, 0>(qr, er, r), pr + (pr - orr[r]));
            ns[r] = inside ? sn : os[r];
            nr[r] = inside ? rn : orr[r];
            if (inside) {
                il[r] = add_nc(il[r], mul_nc(sn, sn));
                if constexpr (MODE == 0) ac[r] = add_nc(ac[r], mul_nc(sn, pr));
                else ac[r] = add_nc(ac[r], mul_nc(pr, mul_nc(c[r], lapS)));
            }
        }
        if (inj_on) {
            if constexpr (staged) inject_staged<VZ, 0>(st, a.idh2, nr, j - j0, zb);
            else inject_receivers<VZ>(a, nr, j, zb, dat, a.nt - 2 - k);
        }
        if (active) {
            const long off = (long)j * ldz + zb;
            vstore<VZ>(sout + off, ns); vstore<VZ>(rout + off, nr);
            vstore<VZ>(acc + off, ac); vstore<VZ>(ill + off, il);
        }
#pragma unroll
        for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int r = 0; r < VZ; ++r) { qs[m][r] = qs[m + 1][r]; qr[m][r] = qr[m + 1][r]; }
#pragma unroll
        for (int r = 0; r < VZ; ++r) { qs[8][r] = psn[r]; qr[8][r] = prn[r]; }
    }
}

// Two reverse steps per launch, tiled like fwd2_kernel, for both wavefields.
// The step-1 image / illumination contributions (rounded products) wait in a
// 5-column queue; image and illumination are read and written once per two
// steps and summed in the same order as two single steps.
template <int VZ, int MODE, bool ST>
__global__ __launch_bounds__(64) void rev2_kernel(RevStep a, int nzseg, int nxchunk, int xw,
                                                  int nslots, int k)
{
    TileId id;
    if (!tile_of(nzseg, nxchunk, nslots, id)) return;
    const Grid g = a.g;
    const int lane = threadIdx.x;
    const int ldz = g.ldz, nxx = g.nxx;
    const int zs = id.zseg * (64 * VZ - 8);
    const int zb = zs + lane * VZ;
    const bool active = zb < ldz;
    const bool own = active && (zb >= zs + 4) && (zb < zs + 64 * VZ - 4);
    const int x0 = 4 + id.xc * xw;
    const int x1 = min(x0 + xw, nxx - 4);

    const long soff = (long)id.slot * g.fstride;
    const float *scur = a.scur + soff;
    const float *sold = a.sold + soff;
    float *sout1 = a.sout1 + soff;
    float *sout2 = a.sout2 + soff;
    const float *rcur = a.rcur + soff;
    const float *rold = a.rold + soff;
    float *rout1 = a.rout1 + soff;
    float *rout2 = a.rout2 + soff;
    const float *cc = a.c + (long)id.slot * a.cstride;
    float *acc = a.acc + soff;
    float *ill = a.ill + soff;
    const bool inj1 = (k + 1 < a.nt);
    const bool inj2 = (k + 2 < a.nt);
    const int drow = a.data_by_shot ? (*a.shot0 + id.slot) : id.slot;
    const float *dat = a.data + (long)drow * a.rec.nrec * a.nt;
    const int xa = x0 - 4;                           // window [x0 - 4, x1 + 4)
    __shared__ RevStage st;
    constexpr bool staged = ST;     // tables fit the LDS (the host checks)
    if constexpr (ST) (void)stage_rev(a, dat, xa, x1 + 4 - xa, zs, zs + 64 * VZ, inj1, a.nt - 2 - k,
                                        inj2, a.nt - 3 - k, false, 0, st, lane);

    float qs1[9][VZ], qr1[9][VZ];    // S_cur, R_cur: columns j-4 .. j+4
    float qs2[9][VZ], qr2[9][VZ];    // step-1 S, R: columns j-8 .. j
    float iq[5][VZ], aq[5][VZ];      // step-1 illum / image contributions
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        ldcol<VZ>(qs1[m], scur, x0 - 8 + m, nxx, ldz, zb);
        ldcol<VZ>(qr1[m], rcur, x0 - 8 + m, nxx, ldz, zb);
        vzero<VZ>(qs2[m]); vzero<VZ>(qr2[m]);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) { vzero<VZ>(iq[m]); vzero<VZ>(aq[m]); }
    float osn[VZ], orn[VZ], cn[VZ], psn[VZ], prn[VZ];
    ldcol<VZ>(osn, sold, x0 - 4, nxx, ldz, zb);
    ldcol<VZ>(orn, rold, x0 - 4, nxx, ldz, zb);
    ldcol<VZ>(cn, cc, x0 - 4, nxx, ldz, zb);
    vzero<VZ>(psn); vzero<VZ>(prn);
    float4 hs = ldhalo<VZ>(scur, x0 - 4, 4, nxx - 4, ldz, zs, lane);
    float4 hr = ldhalo<VZ>(rcur, x0 - 4, 4, nxx - 4, ldz, zs, lane);
    __shared__ float cring[5][64 * VZ];              // c of the last five columns
    int slot = 0;

    for (int j = x0 - 4; j < x1 + 4; ++j) {
        float os[VZ], orr[VZ], cj[VZ];
#pragma unroll
        for (int r = 0; r < VZ; ++r) { os[r] = osn[r]; orr[r] = orn[r]; cj[r] = cn[r]; }
        vstore<VZ>(&cring[slot][lane * VZ], cj);
        slot = (slot == 4) ? 0 : slot + 1;          // = the slot of column j - 4
        const int jj = j - 4;
        if (j + 1 < x1 + 4) {
            ldcol<VZ>(psn, scur, j + 5, nxx, ldz, zb);
            ldcol<VZ>(prn, rcur, j + 5, nxx, ldz, zb);
            ldcol<VZ>(osn, sold, j + 1, nxx, ldz, zb);
            ldcol<VZ>(orn, rold, j + 1, nxx, ldz, zb);
            ldcol<VZ>(cn, cc, j + 1, nxx, ldz, zb);
        }
        // ---- step 1 at column j
        float ns[VZ], nr[VZ], ci[VZ], ca[VZ];
        vzero<VZ>(ns); vzero<VZ>(nr); vzero<VZ>(ci); vzero<VZ>(ca);
        if (j >= 4 && j < nxx - 4) {
            float es[VZ + 8], er[VZ + 8];
            zext<VZ>(es, qs1[4], hs);
            zext<VZ>(er, qr1[4], hr);
            hs = ldhalo<VZ>(scur, j + 1, 4, nxx - 4, ldz, zs, lane);
            hr = ldhalo<VZ>(rcur, j + 1, 4, nxx - 4, ldz, zs, lane);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const bool inside = (i >= 4) && (i < g.nzz - 4);
                const float ps = qs1[4][r], pr = qr1[4][r];
                const float lapS = lap8r<VZ, 0>(qs1, es, r);
                const float sn = fmaf(cj[r], lapS, ps + (ps - os[r]));
                const float rn = fmaf(cj[r], lap8r<VZ, 0>(qr1, er, r), pr + (pr - orr[r]));
                ns[r] = inside ? sn : os[r];
                nr[r] = inside ? rn : orr[r];
                if (inside) {
                    ci[r] = mul_nc(sn, sn);
                    ca[r] = (MODE == 0) ? mul_nc(sn, pr) : mul_nc(pr, mul_nc(cj[r], lapS));
                }
            }
            if (inj1) {
                if constexpr (staged) inject_staged<VZ, 0>(st, a.idh2, nr, j - xa, zb);
                else inject_receivers<VZ>(a, nr, j, zb, dat, a.nt - 2 - k);
            }
            if (own && j >= x0 && j < x1) {
                vstore<VZ>(sout1 + (long)j * ldz + zb, ns);
                vstore<VZ>(rout1 + (long)j * ldz + zb, nr);
            }
        } else {
            hs = ldhalo<VZ>(scur, j + 1, 4, nxx - 4, ldz, zs, lane);
            hr = ldhalo<VZ>(rcur, j + 1, 4, nxx - 4, ldz, zs, lane);
        }
#pragma unroll
        for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int r = 0; r < VZ; ++r) { qs2[m][r] = qs2[m + 1][r]; qr2[m][r] = qr2[m + 1][r]; }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int r = 0; r < VZ; ++r) { iq[m][r] = iq[m + 1][r]; aq[m][r] = aq[m + 1][r]; }
#pragma unroll
        for (int r = 0; r < VZ; ++r) {
            qs2[8][r] = ns[r]; qr2[8][r] = nr[r]; iq[4][r] = ci[r]; aq[4][r] = ca[r];
        }

        // ---- step 2 at column jj
        if (jj >= x0) {
            float ac[VZ], il[VZ], cjj[VZ];
            ldcol<VZ>(ac, acc, jj, nxx, ldz, zb);
            ldcol<VZ>(il, ill, jj, nxx, ldz, zb);
            vload<VZ>(cjj, &cring[slot][lane * VZ]);   // c at jj (parked 4 columns ago)
            float es[VZ + 8], er[VZ + 8];
            zext_local<VZ>(es, qs2[4]);
            zext_local<VZ>(er, qr2[4]);
            float ns2[VZ], nr2[VZ];
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const bool inside = (i >= 4) && (i < g.nzz - 4);
                const float os2 = qs1[0][r];            // S_cur at column jj
                const float or2 = qr1[0][r];            // R_k at column jj
                const float ps = qs2[4][r], pr = qr2[4][r];
                const float c2 = cjj[r];
                const float lapS = lap8r<VZ, 0>(qs2, es, r);
                const float sn = fmaf(c2, lapS, ps + (ps - os2));
                const float rn = fmaf(c2, lap8r<VZ, 0>(qr2, er, r), pr + (pr - or2));
                ns2[r] = inside ? sn : os2;
                nr2[r] = inside ? rn : or2;
                if (inside) {
                    il[r] = add_nc(add_nc(il[r], iq[0][r]), mul_nc(sn, sn));
                    const float c2a = (MODE == 0) ? mul_nc(sn, pr) : mul_nc(pr, mul_nc(c2, lapS));
                    ac[r] = add_nc(add_nc(ac[r], aq[0][r]), c2a);
                }
            }
            if (inj2) {
                if constexpr (staged) inject_staged<VZ, 1>(st, a.idh2, nr2, jj - xa, zb);
                else inject_receivers<VZ>(a, nr2, jj, zb, dat, a.nt - 3 - k);
            }
            if (own) {
                const long off = (long)jj * ldz + zb;
                vstore<VZ>(sout2 + off, ns2);
                vstore<VZ>(rout2 + off, nr2);
                vstore<VZ>(acc + off, ac);
                vstore<VZ>(ill + off, il);
            }
        }
#pragma unroll
        for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int r = 0; r < VZ; ++r) { qs1[m][r] = qs1[m + 1][r]; qr1[m][r] = qr1[m + 1][r]; }
#pragma unroll
        for (int r = 0; r < VZ; ++r) { qs1[8][r] = psn[r]; qr1[8][r] = prn[r]; }
    }
}

// Three reverse steps per launch, tiled like fwd3_kernel, for both
// wavefields.  Image and illumination of a column are read when step 1
// reaches it, take the three steps' rounded products in step order in a
// 9-column queue of partial sums, and are written once when step 3 leaves it.
template <int VZ, int MODE, bool ST>
__global__ __launch_bounds__(64) void rev3_kernel(RevStep a, int nzseg, int nxchunk, int xw,
                                                  int nslots, int k)
{
    TileId id;
    if (!tile_of(nzseg, nxchunk, nslots, id)) return;
    const Grid g = a.g;
    const int lane = threadIdx.x;
    const int ldz = g.ldz, nxx = g.nxx;
    const int zs = id.zseg * (64 * VZ - 16) - 4;    // first computed row; owned from zs + 8
    const int zb = zs + lane * VZ;
    const int zl = zb < 0 ? 0 : zb;                  // load row: rows above 0 feed rim rows only
    const unsigned uzl = unsigned(zl);
    const bool active = zb < ldz;
    const bool own = active && (zb >= zs + 8) && (zb < zs + 64 * VZ - 8);
    const int x0 = 4 + id.xc * xw;
    const int x1 = min(x0 + xw, nxx - 4);
    // odd chunks march right to left (column c of the march is PC(c))
    const bool back = (id.xc & 1) != 0;
    const int mir = x0 + x1 - 1;
    auto PC = [&](int c) { return back ? mir - c : c; };

    const long soff = (long)id.slot * g.fstride;
    const float *scur = a.scur + soff;
    const float *sold = a.sold + soff;
    float *sout1 = a.sout1 + soff;
    float *sout2 = a.sout2 + soff;
    const float *rcur = a.rcur + soff;
    const float *rold = a.rold + soff;
    float *rout1 = a.rout1 + soff;
    float *rout2 = a.rout2 + soff;
    const float *cc = a.c + (long)id.slot * a.cstride;
    float *acc = a.acc + soff;
    float *ill = a.ill + soff;
    const bool inj1 = (k + 1 < a.nt);
    const bool inj2 = (k + 2 < a.nt);
    const bool inj3 = (k + 3 < a.nt);
    const int drow = a.data_by_shot ? (*a.shot0 + id.slot) : id.slot;
    const float *dat = a.data + (long)drow * a.rec.nrec * a.nt;
    const int xa = x0 - 8;                           // window [x0 - 8, x1 + 8)
    __shared__ RevStage st;
    constexpr bool staged = ST;     // tables fit the LDS (the host checks)
    if constexpr (ST) (void)stage_rev(a, dat, xa, x1 + 8 - xa, zs, zs + 64 * VZ, inj1, a.nt - 2 - k,
                                        inj2, a.nt - 3 - k, inj3, a.nt - 4 - k, st, lane);

    float qs1[9][VZ], qr1[9][VZ];    // S_cur, R_cur
    float qs2[9][VZ], qr2[9][VZ];    // step-1 S, R
    float qs3[9][VZ], qr3[9][VZ];    // step-2 S, R
    float pa[9][VZ], pi[9][VZ];      // image / illumination partial sums
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        ldcolc<VZ>(qs1[m], scur, PC(x0 - 12 + m), nxx, ldz, zl);
        ldcolc<VZ>(qr1[m], rcur, PC(x0 - 12 + m), nxx, ldz, zl);
        vzero<VZ>(qs2[m]); vzero<VZ>(qr2[m]);
        vzero<VZ>(qs3[m]); vzero<VZ>(qr3[m]);
        vzero<VZ>(pa[m]); vzero<VZ>(pi[m]);
    }
    // halos (ldhalo1) of the S_cur / R_cur queue columns, in the queues' slots
    float hqs[9], hqr[9];
#pragma unroll
    for (int m = 4; m < 9; ++m) {
        hqs[m] = ldhalo1<VZ>(scur, PC(x0 - 12 + m), nxx, ldz, zs, lane);
        hqr[m] = ldhalo1<VZ>(rcur, PC(x0 - 12 + m), nxx, ldz, zs, lane);
    }
    hqs[0] = hqs[1] = hqs[2] = hqs[3] = 0.f;
    hqr[0] = hqr[1] = hqr[2] = hqr[3] = 0.f;
    float osn[VZ], orn[VZ], cn[VZ], psn[VZ], prn[VZ], an[VZ], in[VZ];
    float hsn, hrn;
    ldcolc<VZ>(osn, sold, PC(x0 - 8), nxx, ldz, zl);
    ldcolc<VZ>(orn, rold, PC(x0 - 8), nxx, ldz, zl);
    ldcolc<VZ>(cn, cc, PC(x0 - 8), nxx, ldz, zl);
    vzero<VZ>(an); vzero<VZ>(in);
    __shared__ float cring[9][64 * VZ];              // c of the last nine columns
    const int jend = x1 + 8;                         // window [x0 - 8, x1 + 8)
    int j = x0 - 8;

    // One column j.  P is the phase of the unrolled loop: slot (m + P) % 9 of
    // a queue holds its column m (q1: j-4+m; q2, pa, pi: j-8+m; q3: j-12+m
    // once the column's results are in; cring: j-8+m), so the queues advance
    // without register moves.  Steps 1 and 2 run on every column: a column
    // outside [4, nxx-4) takes its old value, which is a zero rim value.
    auto column = [&](auto phase) {
        constexpr int P = decltype(phase)::value;
        constexpr int C0 = P % 9, C4 = (P + 4) % 9, C8 = (P + 8) % 9;
        float os[VZ], orr[VZ], cj[VZ];
#pragma unroll
        for (int r = 0; r < VZ; ++r) {
            os[r] = osn[r]; orr[r] = orn[r]; cj[r] = cn[r];
            pa[C8][r] = an[r]; pi[C8][r] = in[r];
        }
        vstore<VZ>(&cring[C8][lane * VZ], cj);
        ldcolc<VZ>(psn, scur, PC(j + 5), nxx, ldz, zl);
        ldcolc<VZ>(prn, rcur, PC(j + 5), nxx, ldz, zl);
        // the halo rows of column j + 5 beside its main load: mostly the
        // same cache lines
        hsn = ldhalo1<VZ>(scur, PC(j + 5), nxx, ldz, zs, lane);
        hrn = ldhalo1<VZ>(rcur, PC(j + 5), nxx, ldz, zs, lane);
        ldcolc<VZ>(osn, sold, PC(j + 1), nxx, ldz, zl);
        ldcolc<VZ>(orn, rold, PC(j + 1), nxx, ldz, zl);
        ldcolc<VZ>(cn, cc, PC(j + 1), nxx, ldz, zl);
        if (j + 1 >= x0 && j + 1 < x1) {             // image / illumination of owned columns
            vload<VZ>(an, (acc + (long)PC(j + 1) * ldz) + uzl);
            vload<VZ>(in, (ill + (long)PC(j + 1) * ldz) + uzl);
        }
        // ---- step 1 at column j
        const int pj = PC(j);
        const bool jin1 = (pj >= 4) && (pj < nxx - 4);
        {
            float es[VZ + 8], er[VZ + 8];
            zext1<VZ>(es, qs1[C4], hqs[C4]);
            zext1<VZ>(er, qr1[C4], hqr[C4]);
            float dzs[VZ + 1], dzr[VZ + 1];
            zdiff<VZ>(dzs, es);
            zdiff<VZ>(dzr, er);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const bool inside = jin1 && (i >= 4) && (i < g.nzz - 4);
                const float ps = qs1[C4][r], pr = qr1[C4][r];
                const float lapS = lap8z<VZ, C0>(qs1, es, dzs, r);
                const float sn = fmaf(cj[r], lapS, ps + (ps - os[r]));
                const float rn = fmaf(cj[r], lap8z<VZ, C0>(qr1, er, dzr, r), pr + (pr - orr[r]));
                qs2[C8][r] = inside ? sn : os[r];
                qr2[C8][r] = inside ? rn : orr[r];
                if (inside) {
                    pi[C8][r] = add_nc(pi[C8][r], mul_nc(sn, sn));
                    pa[C8][r] = add_nc(pa[C8][r], (MODE == 0) ? mul_nc(sn, pr)
                                                              : mul_nc(pr, mul_nc(cj[r], lapS)));
                }
            }
            if (inj1 && jin1) {
                if constexpr (staged) inject_staged<VZ, 0>(st, a.idh2, qr2[C8], pj - xa, zb);
                else inject_receivers<VZ>(a, qr2[C8], pj, zb, dat, a.nt - 2 - k);
            }
        }
        // ---- step 2 at column jj = j - 4, from x0 - 4 on (step 3 reads no
        // column before it; the queue slots of the skipped columns are unused)
        const int jj = j - 4, pjj = PC(jj);
        const bool jin2 = (pjj >= 4) && (pjj < nxx - 4);
        if (jj >= x0 - 4) {
            float es[VZ + 8], er[VZ + 8], c2v[VZ];
            zext_local<VZ>(es, qs2[C4]);
            zext_local<VZ>(er, qr2[C4]);
            vload<VZ>(c2v, &cring[C4][lane * VZ]);
            float dzs[VZ + 1], dzr[VZ + 1];
            zdiff<VZ>(dzs, es);
            zdiff<VZ>(dzr, er);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const bool inside = jin2 && (i >= 4) && (i < g.nzz - 4);
                const float c2 = c2v[r];
                const float os2 = qs1[C0][r];           // S_cur at column jj
                const float or2 = qr1[C0][r];           // R_cur at column jj
                const float ps = qs2[C4][r], pr = qr2[C4][r];
                const float lapS = lap8z<VZ, C0>(qs2, es, dzs, r);
                const float sn = fmaf(c2, lapS, ps + (ps - os2));
                const float rn = fmaf(c2, lap8z<VZ, C0>(qr2, er, dzr, r), pr + (pr - or2));
                qs3[C8][r] = inside ? sn : os2;
                qr3[C8][r] = inside ? rn : or2;
                if (inside) {
                    pi[C4][r] = add_nc(pi[C4][r], mul_nc(sn, sn));
                    pa[C4][r] = add_nc(pa[C4][r], (MODE == 0) ? mul_nc(sn, pr)
                                                              : mul_nc(pr, mul_nc(c2, lapS)));
                }
            }
            if (inj2 && jin2) {
                if constexpr (staged) inject_staged<VZ, 1>(st, a.idh2, qr3[C8], pjj - xa, zb);
                else inject_receivers<VZ>(a, qr3[C8], pjj, zb, dat, a.nt - 3 - k);
            }
            if (own && jj >= x0 && jj < x1) {
                vstore<VZ>((sout1 + (long)pjj * ldz) + uzl, qs3[C8]);
                vstore<VZ>((rout1 + (long)pjj * ldz) + uzl, qr3[C8]);
            }
        }
        // ---- step 3 at column jjj = j - 8 (an owned column)
        const int jjj = j - 8;
        if (jjj >= x0) {
            const int pjjj = PC(jjj);
            float c3[VZ];
            vload<VZ>(c3, &cring[C0][lane * VZ]);
            float es[VZ + 8], er[VZ + 8];
            zext_local<VZ>(es, qs3[C4]);
            zext_local<VZ>(er, qr3[C4]);
            float ns3[VZ], nr3[VZ], fa[VZ], fi[VZ];
            float dzs[VZ + 1], dzr[VZ + 1];
            zdiff<VZ>(dzs, es);
            zdiff<VZ>(dzr, er);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const bool inside = (i >= 4) && (i < g.nzz - 4);
                const float os3 = qs2[C0][r];           // step-1 S at column jjj
                const float or3 = qr2[C0][r];           // step-1 R at column jjj
                const float ps = qs3[C4][r], pr = qr3[C4][r];
                const float lapS = lap8z<VZ, C0>(qs3, es, dzs, r);
                const float sn = fmaf(c3[r], lapS, ps + (ps - os3));
                const float rn = fmaf(c3[r], lap8z<VZ, C0>(qr3, er, dzr, r), pr + (pr - or3));
                ns3[r] = inside ? sn : os3;
                nr3[r] = inside ? rn : or3;
                fi[r] = pi[C0][r];
                fa[r] = pa[C0][r];
                if (inside) {
                    fi[r] = add_nc(fi[r], mul_nc(sn, sn));
                    fa[r] = add_nc(fa[r], (MODE == 0) ? mul_nc(sn, pr)
                                                      : mul_nc(pr, mul_nc(c3[r], lapS)));
                }
            }
            if (inj3) {
                if constexpr (staged) inject_staged<VZ, 2>(st, a.idh2, nr3, pjjj - xa, zb);
                else inject_receivers<VZ>(a, nr3, pjjj, zb, dat, a.nt - 4 - k);
            }
            if (own) {
                const long off = (long)pjjj * ldz;
                vstore<VZ>((sout2 + off) + uzl, ns3);
                vstore<VZ>((rout2 + off) + uzl, nr3);
                vstore<VZ>((acc + off) + uzl, fa);
                vstore<VZ>((ill + off) + uzl, fi);
            }
        }
        // column j + 5 of S_cur, R_cur replaces column j - 4
#pragma unroll
        for (int r = 0; r < VZ; ++r) { qs1[C0][r] = psn[r]; qr1[C0][r] = prn[r]; }
        hqs[C0] = hsn;
        hqr[C0] = hrn;
    };
    for (;;) {
        column(Phase<0>{}); if (++j >= jend) break;
        column(Phase<1>{}); if (++j >= jend) break;
        column(Phase<2>{}); if (++j >= jend) break;
        column(Phase<3>{}); if (++j >= jend) break;
        column(Phase<4>{}); if (++j >= jend) break;
        column(Phase<5>{}); if (++j >= jend) break;
        column(Phase<6>{}); if (++j >= jend) break;
        column(Phase<7>{}); if (++j >= jend) break;
        column(Phase<8>{}); if (++j >= jend) break;
    }
}

// Four reverse steps per launch: the three-step march with a fourth pair of
// register queues (steps on the chunk +- 12, 8, 4, 0 columns and 64*VZ,
// -8, -16, -24 rows), so the S, R, c, image and illumination bytes of four
// steps are those of three.  c of the last thirteen columns waits in LDS.
// The image (or gradient) and illumination products of a column are taken
// two steps at a time from values still in the queues -- steps 1 and 2 when
// step 2 reaches the column, steps 3 and 4 when step 4 does -- and added to
// the column's sums in step order, with the partial sums in LDS in between,
// so the result has the bits of every other launch shape.  The gradient's
// products need c lap S of steps 1 and 3, which wait in register rings.
template <int VZ, int MODE, bool ST>
__global__ __launch_bounds__(64, 2) void rev4_kernel(RevStep a, int nzseg, int nxchunk, int xw,
                                                  int nslots, int k)
{
    TileId id;
    if (!tile_of(nzseg, nxchunk, nslots, id)) return;
    const Grid g = a.g;
    const int lane = threadIdx.x;
    const int ldz = g.ldz, nxx = g.nxx;
    const int zs = id.zseg * (64 * VZ - 24) - 8;    // first computed row; owned from zs + 12
    const int zb = zs + lane * VZ;
    const int zl = zb < 0 ? 0 : zb;
    const unsigned uzl = unsigned(zl);
    const bool active = zb < ldz;
    const bool own = active && (zb >= zs + 12) && (zb < zs + 64 * VZ - 12);
    const int x0 = 4 + id.xc * xw;
    const int x1 = min(x0 + xw, nxx - 4);
    const bool back = (id.xc & 1) != 0;
    const int mir = x0 + x1 - 1;
    auto PC = [&](int c) { return back ? mir - c : c; };

    const long soff = (long)id.slot * g.fstride;
    const float *scur = a.scur + soff;
    const float *sold = a.sold + soff;
    float *sout1 = a.sout1 + soff;
    float *sout2 = a.sout2 + soff;
    const float *rcur = a.rcur + soff;
    const float *rold = a.rold + soff;
    float *rout1 = a.rout1 + soff;
    float *rout2 = a.rout2 + soff;
    const float *cc = a.c + (long)id.slot * a.cstride;
    float *acc = a.acc + soff;
    float *ill = a.ill + soff;
    const bool inj1 = (k + 1 < a.nt), inj2 = (k + 2 < a.nt);
    const bool inj3 = (k + 3 < a.nt), inj4 = (k + 4 < a.nt);
    const int drow = a.data_by_shot ? (*a.shot0 + id.slot) : id.slot;
    const float *dat = a.data + (long)drow * a.rec.nrec * a.nt;
    const int xa = x0 - 12;                          // window [x0 - 12, x1 + 12)
    __shared__ RevStage4 st;
    constexpr bool staged = ST;
    if constexpr (ST) (void)stage_rev4(a, dat, xa, x1 + 12 - xa, zs, zs + 64 * VZ, k, st, lane);

    float qs1[9][VZ], qr1[9][VZ];    // S_cur, R_cur
    float qs2[9][VZ], qr2[9][VZ];    // step-1 S, R
    float qs3[9][VZ], qr3[9][VZ];    // step-2 S, R
    float qs4[9][VZ], qr4[9][VZ];    // step-3 S, R
    // halos (ldhalo1) of the q1 columns; the gradient build, short of
    // registers, loads a float4 halo one column ahead instead
    constexpr int NH = MODE == 0 ? 9 : 1;
    float hqs[NH], hqr[NH];
    float4 h4s, h4r;
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        ldcolc<VZ>(qs1[m], scur, PC(x0 - 16 + m), nxx, ldz, zl);
        ldcolc<VZ>(qr1[m], rcur, PC(x0 - 16 + m), nxx, ldz, zl);
        vzero<VZ>(qs2[m]); vzero<VZ>(qr2[m]);
        vzero<VZ>(qs3[m]); vzero<VZ>(qr3[m]);
        vzero<VZ>(qs4[m]); vzero<VZ>(qr4[m]);
    }
    if constexpr (MODE == 0) {
#pragma unroll
        for (int m = 4; m < 9; ++m) {
            hqs[m] = ldhalo1<VZ>(scur, PC(x0 - 16 + m), nxx, ldz, zs, lane);
            hqr[m] = ldhalo1<VZ>(rcur, PC(x0 - 16 + m), nxx, ldz, zs, lane);
        }
        hqs[0] = hqs[1] = hqs[2] = hqs[3] = 0.f;
        hqr[0] = hqr[1] = hqr[2] = hqr[3] = 0.f;
    } else {
        h4s = ldhaloc<VZ>(scur, PC(x0 - 12), nxx, ldz, zs, lane);
        h4r = ldhaloc<VZ>(rcur, PC(x0 - 12), nxx, ldz, zs, lane);
    }
    float osn[VZ], orn[VZ], cn[VZ], psn[VZ], prn[VZ], an[VZ], in[VZ];
    float hsn, hrn;
    ldcolc<VZ>(osn, sold, PC(x0 - 12), nxx, ldz, zl);
    ldcolc<VZ>(orn, rold, PC(x0 - 12), nxx, ldz, zl);
    ldcolc<VZ>(cn, cc, PC(x0 - 12), nxx, ldz, zl);
    vzero<VZ>(an); vzero<VZ>(in);
    __shared__ float cring[13][64 * VZ];             // c of the last thirteen columns
    // image / illumination partial sums (steps 1-2) of the columns j-12 .. j-4
    // in slot (m + P) % 9: the lane's VZ image rows, then its VZ illumination rows
    __shared__ __align__(16) float pring[9][64 * 2 * VZ];
    // gradient mode: c lap S_0 of columns j-8 .. j (g1q, the q2 slots) and
    // c lap S_2 of columns j-16 .. j-8 (g3q, the q4 slots)
    float g1q[9][MODE == 1 ? VZ : 1], g3q[9][MODE == 1 ? VZ : 1];
    const int jend = x1 + 12;
    int j = x0 - 12;
    int cs = 0;                                      // cring slot of column j

    auto column = [&](auto phase) {
        constexpr int P = decltype(phase)::value;
        constexpr int C0 = P % 9, C4 = (P + 4) % 9, C8 = (P + 8) % 9;
        float os[VZ], orr[VZ], cj[VZ], a2[VZ], i2[VZ];
#pragma unroll
        for (int r = 0; r < VZ; ++r) {
            os[r] = osn[r]; orr[r] = orn[r]; cj[r] = cn[r];
            a2[r] = an[r]; i2[r] = in[r];
        }
        vstore<VZ>(&cring[cs][lane * VZ], cj);
        ldcolc<VZ>(psn, scur, PC(j + 5), nxx, ldz, zl);
        ldcolc<VZ>(prn, rcur, PC(j + 5), nxx, ldz, zl);
        if constexpr (MODE == 0) {
            hsn = ldhalo1<VZ>(scur, PC(j + 5), nxx, ldz, zs, lane);
            hrn = ldhalo1<VZ>(rcur, PC(j + 5), nxx, ldz, zs, lane);
        }
        ldcolc<VZ>(osn, sold, PC(j + 1), nxx, ldz, zl);
        ldcolc<VZ>(orn, rold, PC(j + 1), nxx, ldz, zl);
        ldcolc<VZ>(cn, cc, PC(j + 1), nxx, ldz, zl);
        if (j - 3 >= x0 && j - 3 < x1) {             // sums of the column step 2 reaches next
            vload<VZ>(an, (acc + (long)PC(j - 3) * ldz) + uzl);
            vload<VZ>(in, (ill + (long)PC(j - 3) * ldz) + uzl);
        }
        const int cs2 = cs >= 4 ? cs - 4 : cs + 9;       // cring slot of column j - 4
        const int cs3 = cs >= 8 ? cs - 8 : cs + 5;       // j - 8
        const int cs4 = cs == 12 ? 0 : cs + 1;           // j - 12
        // ---- step 1 at column j
        const int pj = PC(j);
        const bool jin1 = (pj >= 4) && (pj < nxx - 4);
        {
            float es[VZ + 8], er[VZ + 8], dzs[VZ + 1], dzr[VZ + 1];
            if constexpr (MODE == 0) {
                zext1<VZ>(es, qs1[C4], hqs[C4]);
                zext1<VZ>(er, qr1[C4], hqr[C4]);
            } else {
                zext<VZ>(es, qs1[C4], h4s);
                zext<VZ>(er, qr1[C4], h4r);
                h4s = ldhaloc<VZ>(scur, PC(j + 1), nxx, ldz, zs, lane);
                h4r = ldhaloc<VZ>(rcur, PC(j + 1), nxx, ldz, zs, lane);
            }
            zdiff<VZ>(dzs, es);
            zdiff<VZ>(dzr, er);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const bool inside = jin1 && (i >= 4) && (i < g.nzz - 4);
                const float ps = qs1[C4][r], pr = qr1[C4][r];
                const float lapS = lap8z<VZ, C0>(qs1, es, dzs, r);
                const float sn = fmaf(cj[r], lapS, ps + (ps - os[r]));
                const float rn = fmaf(cj[r], lap8z<VZ, C0>(qr1, er, dzr, r), pr + (pr - orr[r]));
                qs2[C8][r] = inside ? sn : os[r];
                qr2[C8][r] = inside ? rn : orr[r];
                if constexpr (MODE == 1) g1q[C8][r] = mul_nc(cj[r], lapS);
            }
            if (inj1 && jin1) {
                if constexpr (staged) inject_staged<VZ, 0>(st, a.idh2, qr2[C8], pj - xa, zb);
                else inject_receivers<VZ>(a, qr2[C8], pj, zb, dat, a.nt - 2 - k);
            }
        }
        // ---- step 2 at column jj = j - 4, from x0 - 8 on; the image and
        // illumination of steps 1 and 2 of the column
        const int jj = j - 4, pjj = PC(jj);
        const bool jin2 = (pjj >= 4) && (pjj < nxx - 4);
        if (jj >= x0 - 8) {
            float es[VZ + 8], er[VZ + 8], c2v[VZ], dzs[VZ + 1], dzr[VZ + 1], pp[2 * VZ];
            zext_local<VZ>(es, qs2[C4]);
            zext_local<VZ>(er, qr2[C4]);
            vload<VZ>(c2v, &cring[cs2][lane * VZ]);
            zdiff<VZ>(dzs, es);
            zdiff<VZ>(dzr, er);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const bool inside = jin2 && (i >= 4) && (i < g.nzz - 4);
                const float os2 = qs1[C0][r], or2 = qr1[C0][r];   // S_cur, R_cur at jj
                const float ps = qs2[C4][r], pr = qr2[C4][r];     // step-1 S, R at jj
                const float lapS = lap8z<VZ, C0>(qs2, es, dzs, r);
                const float sn = fmaf(c2v[r], lapS, ps + (ps - os2));
                const float rn = fmaf(c2v[r], lap8z<VZ, C0>(qr2, er, dzr, r), pr + (pr - or2));
                qs3[C8][r] = inside ? sn : os2;
                qr3[C8][r] = inside ? rn : or2;
                pp[r] = a2[r];
                pp[VZ + r] = i2[r];
                if (inside) {        // step 1 (S_1 R_0 or R_0 c lap S_0), then step 2
                    pp[VZ + r] = add_nc(pp[VZ + r], mul_nc(ps, ps));
                    if constexpr (MODE == 0) pp[r] = add_nc(pp[r], mul_nc(ps, or2));
                    else pp[r] = add_nc(pp[r], mul_nc(or2, g1q[C4][r]));
                    pp[VZ + r] = add_nc(pp[VZ + r], mul_nc(sn, sn));
                    if constexpr (MODE == 0) pp[r] = add_nc(pp[r], mul_nc(sn, pr));
                    else pp[r] = add_nc(pp[r], mul_nc(pr, mul_nc(c2v[r], lapS)));
                }
            }
            vstore<2 * VZ>(&pring[C8][lane * 2 * VZ], pp);
            if (inj2 && jin2) {
                if constexpr (staged) inject_staged<VZ, 1>(st, a.idh2, qr3[C8], pjj - xa, zb);
                else inject_receivers<VZ>(a, qr3[C8], pjj, zb, dat, a.nt - 3 - k);
            }
        }
        // ---- step 3 at column jjj = j - 8, from x0 - 4 on
        const int jjj = j - 8, pjjj = PC(jjj);
        const bool jin3 = (pjjj >= 4) && (pjjj < nxx - 4);
        if (jjj >= x0 - 4) {
            float es[VZ + 8], er[VZ + 8], c3v[VZ], dzs[VZ + 1], dzr[VZ + 1];
            zext_local<VZ>(es, qs3[C4]);
            zext_local<VZ>(er, qr3[C4]);
            vload<VZ>(c3v, &cring[cs3][lane * VZ]);
            zdiff<VZ>(dzs, es);
            zdiff<VZ>(dzr, er);
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const bool inside = jin3 && (i >= 4) && (i < g.nzz - 4);
                const float os3 = qs2[C0][r], or3 = qr2[C0][r];   // step-1 S, R at jjj
                const float ps = qs3[C4][r], pr = qr3[C4][r];
                const float lapS = lap8z<VZ, C0>(qs3, es, dzs, r);
                const float sn = fmaf(c3v[r], lapS, ps + (ps - os3));
                const float rn = fmaf(c3v[r], lap8z<VZ, C0>(qr3, er, dzr, r), pr + (pr - or3));
                qs4[C8][r] = inside ? sn : os3;
                qr4[C8][r] = inside ? rn : or3;
                if constexpr (MODE == 1) g3q[C8][r] = mul_nc(c3v[r], lapS);
            }
            if (inj3 && jin3) {
                if constexpr (staged) inject_staged<VZ, 2>(st, a.idh2, qr4[C8], pjjj - xa, zb);
                else inject_receivers<VZ>(a, qr4[C8], pjjj, zb, dat, a.nt - 4 - k);
            }
            if (own && jjj >= x0 && jjj < x1) {
                vstore<VZ>((sout1 + (long)pjjj * ldz) + uzl, qs4[C8]);
                vstore<VZ>((rout1 + (long)pjjj * ldz) + uzl, qr4[C8]);
            }
        }
        // ---- step 4 at column j4 = j - 12 (an owned column); the image and
        // illumination of steps 3 and 4, added to the column's partial sums
        const int j4 = j - 12;
        if (j4 >= x0) {
            const int pj4 = PC(j4);
            float c4v[VZ], pp[2 * VZ], es[VZ + 8], er[VZ + 8], dzs[VZ + 1], dzr[VZ + 1];
            vload<VZ>(c4v, &cring[cs4][lane * VZ]);
            vload<2 * VZ>(pp, &pring[C0][lane * 2 * VZ]);
            zext_local<VZ>(es, qs4[C4]);
            zext_local<VZ>(er, qr4[C4]);
            zdiff<VZ>(dzs, es);
            zdiff<VZ>(dzr, er);
            float ns4[VZ], nr4[VZ], fa[VZ], fi[VZ];
#pragma unroll
            for (int r = 0; r < VZ; ++r) {
                const int i = zb + r;
                const bool inside = (i >= 4) && (i < g.nzz - 4);
                const float os4 = qs3[C0][r], or4 = qr3[C0][r];   // step-2 S, R at j4
                const float ps = qs4[C4][r], pr = qr4[C4][r];     // step-3 S, R at j4
                const float lapS = lap8z<VZ, C0>(qs4, es, dzs, r);
                const float sn = fmaf(c4v[r], lapS, ps + (ps - os4));
                const float rn = fmaf(c4v[r], lap8z<VZ, C0>(qr4, er, dzr, r), pr + (pr - or4));
                ns4[r] = inside ? sn : os4;
                nr4[r] = inside ? rn : or4;
                fa[r] = pp[r];
                fi[r] = pp[VZ + r];
                if (inside) {        // step 3 (S_3 R_2 or R_2 c lap S_2), then step 4
                    fi[r] = add_nc(fi[r], mul_nc(ps, ps));
                    if constexpr (MODE == 0) fa[r] = add_nc(fa[r], mul_nc(ps, or4));
                    else fa[r] = add_nc(fa[r], mul_nc(or4, g3q[C4][r]));
                    fi[r] = add_nc(fi[r], mul_nc(sn, sn));
                    if constexpr (MODE == 0) fa[r] = add_nc(fa[r], mul_nc(sn, pr));
                    else fa[r] = add_nc(fa[r], mul_nc(pr, mul_nc(c4v[r], lapS)));
                }
            }
            if (inj4) {
                if constexpr (staged) inject_staged<VZ, 3>(st, a.idh2, nr4, pj4 - xa, zb);
                else inject_receivers<VZ>(a, nr4, pj4, zb, dat, a.nt - 5 - k);
            }
            if (own) {
                const long off = (long)pj4 * ldz;
                vstore<VZ>((sout2 + off) + uzl, ns4);
                vstore<VZ>((rout2 + off) + uzl, nr4);
                vstore<VZ>((acc + off) + uzl, fa);
                vstore<VZ>((ill + off) + uzl, fi);
            }
        }
#pragma unroll
        for (int r = 0; r < VZ; ++r) { qs1[C0][r] = psn[r]; qr1[C0][r] = prn[r]; }
        if constexpr (MODE == 0) {
            hqs[C0 % NH] = hsn;
            hqr[C0 % NH] = hrn;
        }
        cs = (cs == 12) ? 0 : cs + 1;
    };
    for (;;) {
        column(Phase<0>{}); if (++j >= jend) break;
        column(Phase<1>{}); if (++j >= jend) break;
        column(Phase<2>{}); if (++j >= jend) break;
        column(Phase<3>{}); if (++j >= jend) break;
        column(Phase<4>{}); if (++j >= jend) break;
        column(Phase<5>{}); if (++j >= jend) break;
        column(Phase<6>{}); if (++j >= jend) break;
        column(Phase<7>{}); if (++j >= jend) break;
        column(Phase<8>{}); if (++j >= jend) break;
    }
}

// Reverse steps on 64 x 64-cell LDS tiles (the reverse counterpart of
// fwdt_kernel, for one or a few shots): S and R of the tile in LDS, c and
// the own cells' image (or gradient) and illumination sums in registers.
// Step k updates S and R on the tile's inner (64 - 8k)^2 cells with the
// column kernels' arithmetic, adds the step's products to the own cells'
// sums in step order (image S_k R_{k-1} or gradient R_{k-1} c lap S_{k-1},
// illumination S_k^2), and then adds the step's data at the receivers of the
// region, one column per thread in receiver-table order (the data of the
// window's first kRevTileRec receivers waits in LDS, the rest is read when
// needed).  Bit-identical to the column kernels.
constexpr int kRevTileRec = 256;

template <int NS, int MODE, int NW>
__global__ __launch_bounds__(64 * NW) void revt_kernel(RevStep a, int ntz, int ntx, int nslots,
                                                       int k)
{
    constexpr int H = 4 * NS, OWN = kTile - 2 * H, CPT = kTile / NW;
    const int tile = xcd_tile(ntz * ntx * nslots);
    if (tile < 0) return;                            // the whole workgroup
    const int tz = tile % ntz, tx = (tile / ntz) % ntx, slot = tile / (ntz * ntx);
    const Grid g = a.g;
    const int ldz = g.ldz, nxx = g.nxx, nzz = g.nzz;
    const int z0 = 4 + tz * OWN - H, x0 = 4 + tx * OWN - H;
    const int z = threadIdx.x & 63, xg = threadIdx.x >> 6;
    const int gz = z0 + z;
    const bool zok = (gz >= 0) & (gz < nzz);
    const long soff = (long)slot * g.fstride;
    const int zlo = 4 + tz * OWN, xlo = 4 + tx * OWN;
    const int zhi = min(zlo + OWN, nzz), xhi = min(xlo + OWN, nxx - 4);
    const bool zown = (gz >= zlo) & (gz < zhi);

    __shared__ float su[2][kTile][kTile], ru[2][kTile][kTile];   // [buffer][column][row]
    __shared__ int scp[kTile - 7];       // receiver offsets of the columns x0+4 .. x0+60
    __shared__ int riz[kRevTileRec];
    __shared__ float rv[NS][kRevTileRec];
    float cr[CPT], ga[CPT], gi[CPT];     // c of the thread's cells; own cells' sums
    {
        float l0[CPT], l1[CPT], l2[CPT], l3[CPT];
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const int x = xg + NW * m, gx = x0 + x;
            const bool ok = zok & (gx >= 0) & (gx < nxx);
            const long idx = soff + (long)gx * ldz + gz;
            const bool need = (x >= 4) & (x < kTile - 4) & (z >= 4) & (z < kTile - 4);
            const bool own = zown & (gx >= xlo) & (gx < xhi);
            l0[m] = ok ? a.scur[idx] : 0.f;
            l1[m] = ok ? a.rcur[idx] : 0.f;
            l2[m] = (ok & need) ? a.sold[idx] : 0.f;
            l3[m] = (ok & need) ? a.rold[idx] : 0.f;
            cr[m] = ok ? a.c[(long)slot * a.cstride + (long)gx * ldz + gz] : 0.f;
            ga[m] = own ? a.acc[idx] : 0.f;
            gi[m] = own ? a.ill[idx] : 0.f;
        }
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const int x = xg + NW * m;
            su[0][x][z] = l0[m];
            ru[0][x][z] = l1[m];
            su[1][x][z] = l2[m];
            ru[1][x][z] = l3[m];
        }
    }
    // receivers of the window columns, one column per thread: offsets and,
    // for the first kRevTileRec, their rows and the launch's data samples
    const bool inj_any = (k + 1 < a.nt);
    const int drow = a.data_by_shot ? (*a.shot0 + slot) : slot;
    const float *dat = a.data + (long)drow * a.rec.nrec * a.nt;
    const int c = threadIdx.x;                       // window column x0 + 4 + c
    const bool colt = inj_any & (c < kTile - 8);
    int e0 = 0;
    if (inj_any) e0 = a.rec.col_ptr[min(max(x0 + 4, 0), nxx)];
    if (colt) {
        const int gx = x0 + 4 + c;
        scp[c] = a.rec.col_ptr[min(max(gx, 0), nxx)] - e0;
        if (c == kTile - 9) scp[c + 1] = a.rec.col_ptr[min(max(gx + 1, 0), nxx)] - e0;
    }
    __syncthreads();
    if (colt) {
        for (int e = scp[c]; e < min(scp[c + 1], kRevTileRec); ++e) {
            riz[e] = a.rec.iz[e0 + e];
            const float *row = dat + (long)a.rec.id[e0 + e] * a.nt;
#pragma unroll
            for (int q = 0; q < NS; ++q) rv[q][e] = (k + 1 + q < a.nt) ? row[a.nt - 2 - k - q] : 0.f;
        }
    }

#pragma unroll
    for (int s = 1; s <= NS; ++s) {
        const float (*cS)[kTile] = su[(s + 1) & 1];      // S_{s-1}
        const float (*cR)[kTile] = ru[(s + 1) & 1];      // R_{s-1}
        float (*nS)[kTile] = su[s & 1];                  // S_{s-2} -> S_s
        float (*nR)[kTile] = ru[s & 1];
        const bool zin = (z >= 4 * s) & (z < kTile - 4 * s);
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const int x = xg + NW * m;                   // the same for the whole wave
            if ((x < 4 * s) | (x >= kTile - 4 * s)) continue;
            const float ps = cS[x][z], pr = cR[x][z];
            const float lapS = lap8_cell(cS, x, z, ps);  // all 64 lanes
            const float lapR = lap8_cell(cR, x, z, pr);
            if (zin) {
                const int gx = x0 + x;
                const float os = nS[x][z], orr = nR[x][z];
                const float sn = fmaf(cr[m], lapS, ps + (ps - os));
                const float rn = fmaf(cr[m], lapR, pr + (pr - orr));
                const bool inside = (gx >= 4) & (gx < nxx - 4) & (gz >= 4) & (gz < nzz - 4);
                const float S = inside ? sn : os;
                nS[x][z] = S;
                nR[x][z] = inside ? rn : orr;
                if (inside & zown & (gx >= xlo) & (gx < xhi)) {
                    gi[m] = add_nc(gi[m], mul_nc(S, S));
                    if constexpr (MODE == 0) ga[m] = add_nc(ga[m], mul_nc(S, pr));
                    else ga[m] = add_nc(ga[m], mul_nc(pr, mul_nc(cr[m], lapS)));
                }
            }
        }
        if (k + s < a.nt) {                              // the step's data at the receivers
            __syncthreads();
            const int lo = 4 * s, hi = kTile - 4 * s;
            if (colt && (c + 4 >= lo) && (c + 4 < hi)) {
                const int x = c + 4, gx = x0 + x;
                if ((gx >= 4) & (gx < nxx - 4))
                    for (int e = scp[c]; e < scp[c + 1]; ++e) {
                        const bool st = e < kRevTileRec;
                        const int zz = (st ? riz[e] : a.rec.iz[e0 + e]) - z0;
                        if (zz < lo || zz >= hi) continue;
                        const float v = st ? rv[s - 1][e]
                                           : dat[(long)a.rec.id[e0 + e] * a.nt + a.nt - 1 - s - k];
                        nR[x][zz] = fmaf(a.idh2, v, nR[x][zz]);
                    }
            }
        }
        __syncthreads();
        float *os_ = (s == NS) ? (NS == 1 ? a.sout1 : a.sout2) : (s == NS - 1 ? a.sout1 : nullptr);
        float *or_ = (s == NS) ? (NS == 1 ? a.rout1 : a.rout2) : (s == NS - 1 ? a.rout1 : nullptr);
        if (os_ && zown) {
#pragma unroll
            for (int m = 0; m < CPT; ++m) {
                const int x = xg + NW * m, gx = x0 + x;
                if ((gx >= xlo) & (gx < xhi)) {
                    const long idx = soff + (long)gx * ldz + gz;
                    os_[idx] = nS[x][z];
                    or_[idx] = nR[x][z];
                }
            }
        }
    }
    if (zown) {
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const int gx = x0 + xg + NW * m;
            if ((gx >= xlo) & (gx < xhi)) {
                const long idx = soff + (long)gx * ldz + gz;
                a.acc[idx] = ga[m];
                a.ill[idx] = gi[m];
            }
        }
    }
}

// ---------------------------------------------------------------- helpers

__global__ void fwd_init_kernel(FwdStep a, float *field, int nslots)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslots || a.nt < 1) return;
    const int gshot = *a.shot0 + slot;
    const long idx = (long)slot * a.g.fstride + (long)a.src_ix[gshot] * a.g.ldz + a.src_iz[gshot];
    field[idx] += a.wav[0];
}

__global__ void rev_init_kernel(RevStep a, float *field, int nslots)
{
    // one thread per slot: receivers sharing a cell are summed in order
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslots || a.nt < 1) return;
    const int drow = a.data_by_shot ? (*a.shot0 + slot) : slot;
    const float *dat = a.data + (long)drow * a.rec.nrec * a.nt;
    float *ra = field + (long)slot * a.g.fstride;
    for (int j = 0; j < a.g.nxx; ++j)
        for (int m = a.rec.col_ptr[j]; m < a.rec.col_ptr[j + 1]; ++m)
            ra[(long)j * a.g.ldz + a.rec.iz[m]] += a.idh2 * dat[(long)a.rec.id[m] * a.nt + a.nt - 1];
}

__global__ void fill_kernel(float *p, long n, float v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void set_int_kernel(int *p, int v) { *p = v; }

__global__ void pad_model_kernel(const float *vp, int nz, int nx, Grid g, float *out)
{
    const long n = (long)g.ldz * g.nxx;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const int i = (int)(idx % g.ldz), j = (int)(idx / g.ldz);
        float v = 0.f;
        if (i < g.nzz) {
            const int ii = min(max(i - g.nb, 0), nz - 1);
            const int jj = min(max(j - g.nb, 0), nx - 1);
            v = vp[(long)jj * nz + ii];
        }
        out[idx] = v;
    }
}

__global__ void coef_kernel(const float *v, long n, float k, float *c)
{
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const float x = v[idx];
        c[idx] = (x * x) * k;
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ULL;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z;
}

__device__ __forceinline__ float rb_draw(uint64_t key, uint64_t cell, int k)
{
#pragma clang fp contract(off)
    const uint64_t h = mix64(key + (2 * cell + (uint64_t)k + 1) * kGolden);
    const int32_t s = (int32_t)(h >> 40) - 8388608;
    return (float)s * (1.0f / 8388608.0f);
}

// Seeded random boundary: pad cells take the edge value plus a ramped
// uniform perturbation, once for the x pad and once for the z pad, clamped to
// [vlo, vhi] (the structure of random_boundary_bg, modeling.cu:487-524, with
// a counter-based draw).  FMA contraction is off in this kernel and every
// float operation is one correctly rounded IEEE operation (the division
// included), so the result is bit-identical to a sequential C statement of
// the same generator.
__global__ void rb_kernel(const float *vpad, Grid g, int nslots, const int *shot0,
                          uint32_t iteration, uint64_t seed, float amp_abs, float amp_rel,
                          float vlo, float vhi, float kc, float *c, float *v_out)
{
#pragma clang fp contract(off)
    const long ncell = (long)g.nzz * g.nxx;
    const int slot = blockIdx.y;
    const uint64_t stream = ((uint64_t)iteration << 32) | (uint32_t)(*shot0 + slot);
    const uint64_t key = mix64(seed ^ mix64(stream + kGolden));
    const float fnb = (float)g.nb;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < ncell;
         idx += (long)gridDim.x * blockDim.x) {
        const int i = (int)(idx % g.nzz), j = (int)(idx / g.nzz);
        const float base = vpad[(long)j * g.ldz + i];
        const bool xpad = (j < g.nb) || (j >= g.nxx - g.nb);
        const bool zpad = (i < g.nb) || (i >= g.nzz - g.nb);
        float val = base;
        if (xpad || zpad) {
            const float ra = amp_rel * base;
            const float A = amp_abs + ra;
            const uint64_t cell = (uint64_t)idx;
            if (xpad) {
                const int dj = (j < g.nb) ? j : (g.nxx - 1 - j);
                const float q = (float)dj / fnb;
                const float f = 1.0f - q;
                const float fa = f * A;
                const float d = fa * rb_draw(key, cell, 0);
                val = val + d;
            }
            if (zpad) {
                const int di = (i < g.nb) ? i : (g.nzz - 1 - i);
                const float q = (float)di / fnb;
                const float f = 1.0f - q;
                const float fa = f * A;
                const float d = fa * rb_draw(key, cell, 1);
                val = val + d;
            }
            val = fminf(fmaxf(val, vlo), vhi);
        }
        if (c) c[(long)slot * g.fstride + (long)j * g.ldz + i] = (val * val) * kc;
        if (v_out) v_out[(long)slot * ncell + idx] = val;
    }
}

// Standard normal draw of pad cell `cell` for ramp k (0: x pad, 1: z pad):
// the Irwin-Hall sum of twelve 16-bit uniforms, u = (v + 1/2) / 2^16, from
// three hashes, minus 6.  Integer sum, then one exact scaling: the same bits
// on any IEEE machine.  Mean 0, variance 1, bounded by +-6.
__device__ __forceinline__ float rb_normal(uint64_t key, uint64_t cell, int k)
{
    uint32_t sum = 0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const uint64_t h = mix64(key + (6 * cell + 3 * (uint64_t)k + (uint64_t)m + 1) * kGolden);
        sum += (uint32_t)(h & 0xFFFF) + (uint32_t)((h >> 16) & 0xFFFF) +
               (uint32_t)((h >> 32) & 0xFFFF) + (uint32_t)(h >> 48);
    }
    return (float)((int32_t)(2 * sum) + 12 - 786432) * (1.0f / 131072.0f);
}

// e^-q for q >= 0 with one rounding per operation: n = (int)(q log2 e),
// y = q - n ln 2 (Cody-Waite: ln 2 split into a short head and a tail, two
// fmaf), e^-q = 2^-n e^-y, the second factor by its Taylor polynomial of
// degree 9 (truncation below 1.1e-8 on [0, ln 2)), evaluated by Horner with
// fmaf.
__device__ __forceinline__ float rb_expneg(float q)
{
#pragma clang fp contract(off)
    const int n = (int)(q * 1.44269504f);
    const float fn = (float)n;
    float y = fmaf(-fn, 0.693145752f, q);      // q - n ln 2, ln 2 split in two
    y = fmaf(-fn, 1.42860682e-6f, y);
    float p = -2.75573192e-6f;
    p = fmaf(p, y, 2.48015873e-5f);
    p = fmaf(p, y, -1.98412698e-4f);
    p = fmaf(p, y, 1.38888889e-3f);
    p = fmaf(p, y, -8.33333333e-3f);
    p = fmaf(p, y, 4.16666667e-2f);
    p = fmaf(p, y, -1.66666667e-1f);
    p = fmaf(p, y, 0.5f);
    p = fmaf(p, y, -1.0f);
    p = fmaf(p, y, 1.0f);
    return ldexpf(p, -n);
}

// The reference's random boundary with seeded draws (FWI_RBC_GAUSSIAN;
// random_boundary_bg, modeling.cu:487-524, and random_boundary_gp,
// modeling.cu:467-485).  Pad cells only:
//   val = base [+ (fx sigma) N0 if x pad] [+ (fz sigma) N1 if z pad]
//   then, over the 3 x 3 blob bins around the cell (x-major, z inner) and
//   the blobs of each bin in table order, for d^2 < 16 r^2:
//   val += a * e^-(d^2 * 0.5 / r^2)
//   val = min(max(val, vlo), vhi)
// with x = j dh, z = i dh and the ramps of rb_kernel.  FMA contraction is
// off; every operation is one correctly rounded IEEE operation, so a
// sequential C statement of the same steps gives the same bits.
__global__ void rbg_kernel(const float *vpad, Grid g, int nslots, const int *shot0,
                           uint32_t iteration, uint64_t seed, float sigma, float vlo, float vhi,
                           float kc, float dh, const RbBlob *blob, const int32_t *bstart,
                           int nbx, int nbz, float bin, float *c, float *v_out)
{
#pragma clang fp contract(off)
    const long ncell = (long)g.nzz * g.nxx;
    const int slot = blockIdx.y;
    const uint64_t stream = ((uint64_t)iteration << 32) | (uint32_t)(*shot0 + slot);
    const uint64_t key = mix64(seed ^ mix64(stream + kGolden));
    const float fnb = (float)g.nb;
    const int nbins = nbx * nbz;
    const int32_t *st = bstart + (long)slot * (nbins + 1);
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < ncell;
         idx += (long)gridDim.x * blockDim.x) {
        const int i = (int)(idx % g.nzz), j = (int)(idx / g.nzz);
        const float base = vpad[(long)j * g.ldz + i];
        const bool xpad = (j < g.nb) || (j >= g.nxx - g.nb);
        const bool zpad = (i < g.nb) || (i >= g.nzz - g.nb);
        float val = base;
        if (xpad || zpad) {
            const uint64_t cell = (uint64_t)idx;
            if (xpad) {
                const int dj = (j < g.nb) ? j : (g.nxx - 1 - j);
                const float q = (float)dj / fnb;
                const float f = 1.0f - q;
                const float fs = f * sigma;
                const float d = fs * rb_normal(key, cell, 0);
                val = val + d;
            }
            if (zpad) {
                const int di = (i < g.nb) ? i : (g.nzz - 1 - i);
                const float q = (float)di / fnb;
                const float f = 1.0f - q;
                const float fs = f * sigma;
                const float d = fs * rb_normal(key, cell, 1);
                val = val + d;
            }
            if (nbins > 0) {
                const float x = (float)j * dh, z = (float)i * dh;
                const int cbx = min((int)(x / bin), nbx - 1);
                const int cbz = min((int)(z / bin), nbz - 1);
                for (int bx = max(cbx - 1, 0); bx <= min(cbx + 1, nbx - 1); ++bx)
                    for (int bz = max(cbz - 1, 0); bz <= min(cbz + 1, nbz - 1); ++bz) {
                        const int b = bz + bx * nbz;
                        for (int k = st[b]; k < st[b + 1]; ++k) {
                            const RbBlob B = blob[k];
                            const float dx = x - B.x;
                            const float dz = z - B.z;
                            const float dx2 = dx * dx;
                            const float dz2 = dz * dz;
                            const float d2 = dx2 + dz2;
                            if (d2 < B.c2) {
                                const float e = rb_expneg(d2 * B.h);
                                const float add = B.a * e;
                                val = val + add;
                            }
                        }
                    }
            }
            val = fminf(fmaxf(val, vlo), vhi);
        }
        if (c) c[(long)slot * g.fstride + (long)j * g.ldz + i] = (val * val) * kc;
        if (v_out) v_out[(long)slot * ncell + idx] = val;
    }
}

__global__ __launch_bounds__(256) void residual_kernel(const float *obs, float *cal,
                                                       const int *shot0, long per_shot,
                                                       long n, double *partials)
{
    __shared__ double red[256];
    const long base_obs = (long)(*shot0) * per_shot;
    double s = 0.0;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x) {
        const float r = obs[base_obs + e] - cal[e];
        cal[e] = r;
        s += (double)r * (double)r;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void sum_partials_kernel(const double *p, int n, double *out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int e = threadIdx.x; e < n; e += 256) s += p[e];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

__global__ void reduce_slots_kernel(float *src, long fstride, long ncell, int nslots, float *dst)
{
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < ncell;
         idx += (long)gridDim.x * blockDim.x) {
        float s = dst[idx];
        for (int k = 0; k < nslots; ++k) {
            s += src[(long)k * fstride + idx];
            src[(long)k * fstride + idx] = 0.f;
        }
        dst[idx] = s;
    }
}

__global__ void crop_kernel(const float *src, Grid g, int nz, int nx, float *dst)
{
    const long n = (long)nz * nx;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const int i = (int)(idx % nz), j = (int)(idx / nz);
        dst[idx] = src[(long)(j + g.nb) * g.ldz + (i + g.nb)];
    }
}

__global__ void copy_kernel(const float4 *src, float4 *dst, long n4)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void zero_sample0_kernel(float *seis, int nrec, int nt, long n)
{
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x)
        seis[idx * nt] = 0.f;
}

inline int grid1d(long n, int block, int cap = 8192)
{
    long b = (n + block - 1) / block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

template <int VZ, bool ABC, bool ST, bool REC = true>
void fwd_launch(const FwdStep &a, const Tiling &tl, int nslots, int t, int steps, int nblocks,
                hipStream_t s)
{
    if (steps == 4)
        hipLaunchKernelGGL((fwd4_kernel<VZ, ABC, ST, REC>), dim3(nblocks), dim3(64), 0, s, a,
                           tl.nzseg, tl.nxchunk, tl.xw, nslots, t);
    else if (steps == 3)
        hipLaunchKernelGGL((fwd3_kernel<VZ, ABC, ST, REC>), dim3(nblocks), dim3(64), 0, s, a,
                           tl.nzseg, tl.nxchunk, tl.xw, nslots, t);
    else if (steps == 2)
        hipLaunchKernelGGL((fwd2_kernel<VZ, ABC, ST, REC>), dim3(nblocks), dim3(64), 0, s, a,
                           tl.nzseg, tl.nxchunk, tl.xw, nslots, t);
    else
        hipLaunchKernelGGL((fwd1_kernel<VZ, ABC, ST, REC>), dim3(nblocks), dim3(64), 0, s, a,
                           tl.nzseg, tl.nxchunk, tl.xw, nslots, t);
}

template <int VZ>
hipError_t fwd_dispatch(const FwdStep &a, const Tiling &tl, int nslots, int t, int steps,
                        hipStream_t s)
{
    const int ntiles = tl.nzseg * tl.nxchunk * nslots;
    const int nblocks = ((ntiles + 7) >> 3) << 3;
    if (a.abc) {
        if (tl.staged) fwd_launch<VZ, true, true>(a, tl, nslots, t, steps, nblocks, s);
        else fwd_launch<VZ, true, false>(a, tl, nslots, t, steps, nblocks, s);
    } else if (a.seis == nullptr) {
        // random-boundary passes that record nothing (migration, gradient):
        // the build without the sampling code, which leaves the kernel
        // fewer scalar registers to spill (the only staged table is unused)
        fwd_launch<VZ, false, true, false>(a, tl, nslots, t, steps, nblocks, s);
    } else {
        if (tl.staged) fwd_launch<VZ, false, true>(a, tl, nslots, t, steps, nblocks, s);
        else fwd_launch<VZ, false, false>(a, tl, nslots, t, steps, nblocks, s);
    }
    return hipGetLastError();
}

template <int NS, int NW>
void fwdt_launch(const FwdStep &a, const Tiling &tl, int nslots, int t, int nblocks,
                 hipStream_t s)
{
    if (a.abc)
        hipLaunchKernelGGL((fwdt_kernel<NS, true, NW>), dim3(nblocks), dim3(64 * NW), 0, s, a,
                           tl.nzseg, tl.nxchunk, nslots, t);
    else
        hipLaunchKernelGGL((fwdt_kernel<NS, false, NW>), dim3(nblocks), dim3(64 * NW), 0, s, a,
                           tl.nzseg, tl.nxchunk, nslots, t);
}

template <int NW>
hipError_t fwdt_steps(const FwdStep &a, const Tiling &tl, int nslots, int t, int steps,
                      int nblocks, hipStream_t s)
{
    switch (steps) {
    case 1: fwdt_launch<1, NW>(a, tl, nslots, t, nblocks, s); break;
    case 2: fwdt_launch<2, NW>(a, tl, nslots, t, nblocks, s); break;
    case 3: fwdt_launch<3, NW>(a, tl, nslots, t, nblocks, s); break;
    case 4: fwdt_launch<4, NW>(a, tl, nslots, t, nblocks, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// xw is the columns of the tile each wave updates: 16, 8 or 4 (4, 8 or 16
// waves per tile)
hipError_t fwdt_dispatch(const FwdStep &a, const Tiling &tl, int nslots, int t, int steps,
                         hipStream_t s)
{
    const int ntiles = tl.nzseg * tl.nxchunk * nslots;
    const int nblocks = ((ntiles + 7) >> 3) << 3;
    switch (tl.xw) {
    case 16: return fwdt_steps<4>(a, tl, nslots, t, steps, nblocks, s);
    case 8: return fwdt_steps<8>(a, tl, nslots, t, steps, nblocks, s);
    case 4: return fwdt_steps<16>(a, tl, nslots, t, steps, nblocks, s);
    default: return hipErrorInvalidValue;
    }
}

template <int VZ, int MODE, bool ST>
hipError_t rev_launch(const RevStep &a, const Tiling &tl, int nslots, int k, int steps,
                      int nblocks, hipStream_t s)
{
    if (steps == 4) {
        // one or two rows per lane (the host checks)
        if constexpr (VZ == 4) return hipErrorInvalidValue;
        else
            hipLaunchKernelGGL((rev4_kernel<VZ, MODE, ST>), dim3(nblocks), dim3(64), 0, s, a,
                               tl.nzseg, tl.nxchunk, tl.xw, nslots, k);
    } else if (steps == 3) {
        // six register queues per lane: four rows per lane do not fit
        if constexpr (VZ == 4) return hipErrorInvalidValue;
        else
            hipLaunchKernelGGL((rev3_kernel<VZ, MODE, ST>), dim3(nblocks), dim3(64), 0, s, a,
                               tl.nzseg, tl.nxchunk, tl.xw, nslots, k);
    } else if (steps == 2) {
        hipLaunchKernelGGL((rev2_kernel<VZ, MODE, ST>), dim3(nblocks), dim3(64), 0, s, a,
                           tl.nzseg, tl.nxchunk, tl.xw, nslots, k);
    } else {
        hipLaunchKernelGGL((rev1_kernel<VZ, MODE, ST>), dim3(nblocks), dim3(64), 0, s, a,
                           tl.nzseg, tl.nxchunk, tl.xw, nslots, k);
    }
    return hipGetLastError();
}

template <int VZ>
hipError_t rev_dispatch(const RevStep &a, const Tiling &tl, int nslots, int k, int steps,
                        hipStream_t s)
{
    const int ntiles = tl.nzseg * tl.nxchunk * nslots;
    const int nblocks = ((ntiles + 7) >> 3) << 3;
    if (a.mode == 0)
        return tl.staged ? rev_launch<VZ, 0, true>(a, tl, nslots, k, steps, nblocks, s)
                         : rev_launch<VZ, 0, false>(a, tl, nslots, k, steps, nblocks, s);
    return tl.staged ? rev_launch<VZ, 1, true>(a, tl, nslots, k, steps, nblocks, s)
                     : rev_launch<VZ, 1, false>(a, tl, nslots, k, steps, nblocks, s);
}

template <int NS, int NW>
void revt_launch(const RevStep &a, const Tiling &tl, int nslots, int k, int nblocks,
                 hipStream_t s)
{
    if (a.mode == 0)
        hipLaunchKernelGGL((revt_kernel<NS, 0, NW>), dim3(nblocks), dim3(64 * NW), 0, s, a,
                           tl.nzseg, tl.nxchunk, nslots, k);
    else
        hipLaunchKernelGGL((revt_kernel<NS, 1, NW>), dim3(nblocks), dim3(64 * NW), 0, s, a,
                           tl.nzseg, tl.nxchunk, nslots, k);
}

template <int NW>
hipError_t revt_steps(const RevStep &a, const Tiling &tl, int nslots, int k, int steps,
                      int nblocks, hipStream_t s)
{
    switch (steps) {
    case 1: revt_launch<1, NW>(a, tl, nslots, k, nblocks, s); break;
    case 2: revt_launch<2, NW>(a, tl, nslots, k, nblocks, s); break;
    case 3: revt_launch<3, NW>(a, tl, nslots, k, nblocks, s); break;
    case 4: revt_launch<4, NW>(a, tl, nslots, k, nblocks, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t revt_dispatch(const RevStep &a, const Tiling &tl, int nslots, int k, int steps,
                         hipStream_t s)
{
    const int ntiles = tl.nzseg * tl.nxchunk * nslots;
    const int nblocks = ((ntiles + 7) >> 3) << 3;
    switch (tl.xw) {
    case 16: return revt_steps<4>(a, tl, nslots, k, steps, nblocks, s);
    case 8: return revt_steps<8>(a, tl, nslots, k, steps, nblocks, s);
    case 4: return revt_steps<16>(a, tl, nslots, k, steps, nblocks, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

int zsegments(int nzz, int vz, int steps)
{
    if (vz == kTileVz) return tile_count(nzz, steps);
    // waves tile the updated rows 4 .. nzz-5; the rim rows stay zero
    const int rows = 64 * vz - 8 * (steps - 1);
    const int n = (nzz - 8 + rows - 1) / rows;
    return n > 1 ? n : 1;
}

int xchunks(int nxx, int vz, int xw, int steps)
{
    if (vz == kTileVz) return tile_count(nxx, steps);
    return (nxx - 8 + xw - 1) / xw;
}

hipError_t launch_fwd_step(const FwdStep &a, const Tiling &tl, int nslots, int t, int steps,
                           hipStream_t s)
{
    switch (tl.vz) {
    case 1: return fwd_dispatch<1>(a, tl, nslots, t, steps, s);
    case 2: return fwd_dispatch<2>(a, tl, nslots, t, steps, s);
    case 4: return fwd_dispatch<4>(a, tl, nslots, t, steps, s);
    case kTileVz: return fwdt_dispatch(a, tl, nslots, t, steps, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rev_step(const RevStep &a, const Tiling &tl, int nslots, int k, int steps,
                           hipStream_t s)
{
    switch (tl.vz) {
    case 1: return rev_dispatch<1>(a, tl, nslots, k, steps, s);
    case 2: return rev_dispatch<2>(a, tl, nslots, k, steps, s);
    case 4: return rev_dispatch<4>(a, tl, nslots, k, steps, s);
    case kTileVz: return revt_dispatch(a, tl, nslots, k, steps, s);
    default: return hipErrorInvalidValue;
    }
}

namespace {
template <class K>
int resident_of(K kernel)
{
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return per_cu * cus;
}

template <int VZ>
int resident_vz(int kind, int steps)
{
    if (kind == 0)
        return steps == 4   ? resident_of(fwd4_kernel<VZ, false, true, false>)
               : steps == 3 ? resident_of(fwd3_kernel<VZ, false, true, false>)
               : steps == 2 ? resident_of(fwd2_kernel<VZ, false, true, false>)
                            : resident_of(fwd1_kernel<VZ, false, true, false>);
    if (kind == 2)
        return steps == 4   ? resident_of(fwd4_kernel<VZ, true, true>)
               : steps == 3 ? resident_of(fwd3_kernel<VZ, true, true>)
               : steps == 2 ? resident_of(fwd2_kernel<VZ, true, true>)
                            : resident_of(fwd1_kernel<VZ, true, true>);
    if constexpr (VZ != 4) {
        if (steps == 4) return resident_of(rev4_kernel<VZ, 0, true>);
        if (steps == 3) return resident_of(rev3_kernel<VZ, 0, true>);
    }
    return steps == 2 ? resident_of(rev2_kernel<VZ, 0, true>)
                      : resident_of(rev1_kernel<VZ, 0, true>);
}
}  // namespace

int resident_waves(int kind, int steps, int vz)
{
    switch (vz) {
    case 1: return resident_vz<1>(kind, steps);
    case 2: return resident_vz<2>(kind, steps);
    case 4: return resident_vz<4>(kind, steps);
    default: return 0;
    }
}

hipError_t launch_fwd_init(const FwdStep &a, float *field, int nslots, hipStream_t s)
{
    hipLaunchKernelGGL(fwd_init_kernel, dim3((nslots + 63) / 64), dim3(64), 0, s, a, field,
                       nslots);
    return hipGetLastError();
}

hipError_t launch_rev_init(const RevStep &a, float *field, int nslots, hipStream_t s)
{
    hipLaunchKernelGGL(rev_init_kernel, dim3((nslots + 63) / 64), dim3(64), 0, s, a, field,
                       nslots);
    return hipGetLastError();
}

hipError_t launch_fill(float *p, long n, float v, hipStream_t s)
{
    hipLaunchKernelGGL(fill_kernel, dim3(grid1d(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_set_int(int *p, int v, hipStream_t s)
{
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, p, v);
    return hipGetLastError();
}

hipError_t launch_pad_model(const float *vp, int nz, int nx, const Grid &g, float *out,
                            hipStream_t s)
{
    hipLaunchKernelGGL(pad_model_kernel, dim3(grid1d((long)g.ldz * g.nxx, 256)), dim3(256), 0,
                       s, vp, nz, nx, g, out);
    return hipGetLastError();
}

hipError_t launch_coef(const float *v, const Grid &g, float k, float *c, hipStream_t s)
{
    const long n = (long)g.ldz * g.nxx;
    hipLaunchKernelGGL(coef_kernel, dim3(grid1d(n, 256)), dim3(256), 0, s, v, n, k, c);
    return hipGetLastError();
}

hipError_t launch_rb(const float *vpad, const Grid &g, int nslots, const int *shot0,
                     uint32_t iteration, uint64_t seed, float amp_abs, float amp_rel,
                     float vlo, float vhi, float k, float *c, float *v_out, hipStream_t s)
{
    const long n = (long)g.nzz * g.nxx;
    hipLaunchKernelGGL(rb_kernel, dim3(grid1d(n, 256, 2048), nslots), dim3(256), 0, s, vpad, g,
                       nslots, shot0, iteration, seed, amp_abs, amp_rel, vlo, vhi, k, c, v_out);
    return hipGetLastError();
}

hipError_t launch_rbg(const float *vpad, const Grid &g, int nslots, const int *shot0,
                      uint32_t iteration, uint64_t seed, float sigma, float vlo, float vhi,
                      float k, float dh, const RbBlob *blob, const int32_t *bstart, int nbx,
                      int nbz, float bin, float *c, float *v_out, hipStream_t s)
{
    const long n = (long)g.nzz * g.nxx;
    hipLaunchKernelGGL(rbg_kernel, dim3(grid1d(n, 256, 2048), nslots), dim3(256), 0, s, vpad, g,
                       nslots, shot0, iteration, seed, sigma, vlo, vhi, k, dh, blob, bstart, nbx,
                       nbz, bin, c, v_out);
    return hipGetLastError();
}

hipError_t launch_residual(const float *obs, float *cal, const int *shot0, long per_shot,
                           int nslots, double *partials, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(residual_kernel, dim3(nblocks), dim3(256), 0, s, obs, cal, shot0,
                       per_shot, per_shot * nslots, partials);
    return hipGetLastError();
}

hipError_t launch_sum_partials(const double *partials, int n, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, partials, n, out);
    return hipGetLastError();
}

hipError_t launch_reduce_slots(float *src, long fstride, long ncell, int nslots, float *dst,
                               hipStream_t s)
{
    hipLaunchKernelGGL(reduce_slots_kernel, dim3(grid1d(ncell, 256)), dim3(256), 0, s, src,
                       fstride, ncell, nslots, dst);
    return hipGetLastError();
}

hipError_t launch_crop(const float *src, const Grid &g, int nz, int nx, float *dst,
                       hipStream_t s)
{
    hipLaunchKernelGGL(crop_kernel, dim3(grid1d((long)nz * nx, 256)), dim3(256), 0, s, src, g,
                       nz, nx, dst);
    return hipGetLastError();
}

hipError_t launch_copy(const float *src, float *dst, long n, hipStream_t s)
{
    const long n4 = n / 4;
    hipLaunchKernelGGL(copy_kernel, dim3(grid1d(n4, 256, 65536)), dim3(256), 0, s,
                       reinterpret_cast<const float4 *>(src), reinterpret_cast<float4 *>(dst), n4);
    return hipGetLastError();
}

hipError_t launch_zero_sample0(float *seis, int nrec, int nt, int nslots, hipStream_t s)
{
    const long n = (long)nrec * nslots;
    hipLaunchKernelGGL(zero_sample0_kernel, dim3(grid1d(n, 256)), dim3(256), 0, s, seis, nrec,
                       nt, n);
    return hipGetLastError();
}

}  // namespace fwi
