// rbc.h -- host side of the reference-style random boundary (FWI_RBC_GAUSSIAN):
// the Poisson-disk blob list of one (seed, stream) and the binned table the
// generator kernel reads.  Host-only (no HIP); compiled without FMA
// contraction so that every float operation is one IEEE operation and a plain
// C statement of the same steps gives the same bits.
//
// Reference: Modeling::set_random_boundary (src/modeling/modeling.cu:431-465),
// random_boundary_gp (modeling.cu:467-485) and poissonDiskSampling
// (src/admin/admin.cpp:115-206): Bridson sampling with minimum distance
// `ratio` over the padded model, the points in the outer half of the pad
// kept, each a Gaussian bump of radius U(0.2, 1) ratio and amplitude
// U(0.5, 1) varVp, positive with probability 1/3.  Here the draws come from
// a seeded counter-based generator, so a stream always gives the same blobs.
//
// The sampler, step by step (all float arithmetic in single precision, one
// rounding per operation):
//   key = mix64(seed ^ mix64(stream + G)); state = mix64(key ^ K);
//   u() : state += G; return (mix64(state) >> 40) * 2^-24        in [0, 1)
//   ring: points (x, z) of [0, x_max) x [0, z_max) that are not inside
//         (h, x_max - h) x (h, z_max - h), h = pad / 2;
//   first point: (u() x_max, u() z_max), drawn again until it is in the ring
//         (at most 10000 draws, else no blobs);
//   then Bridson's loop (Bridson, "Fast Poisson disk sampling in arbitrary
//   dimensions", SIGGRAPH 2007 sketches): pick a = (int)(u() n_active)
//   (capped at n_active - 1); up to 30 candidates, each offset
//   ((2u() - 1) 2r, (2u() - 1) 2r) drawn again until r^2 <= d^2 < 4 r^2;
//   the first candidate inside the domain and the ring with no point closer
//   than r (d^2 < r^2) is accepted; with none, active[a] is replaced by the
//   last active point.  Grid cells of side r * 0.70710678f hold at most one
//   point; the search covers the 5 x 5 cells around a candidate.
//   Per point in sample order: radius r * (0.2 + 0.8 u()), amplitude
//   amp * (0.5 + 0.5 u()), negative unless u() < 1/3.
// G = 0x9E3779B97F4A7C15, K = 0xD1B54A32D192ED03, mix64 the SplitMix64
// finaliser (Steele, Lea, Flood, OOPSLA 2014).
#pragma once
#include <cstdint>
#include <vector>

namespace fwi {

struct RbBlob {
    float x, z;     // centre [m] on the padded grid (x = j dh, z = i dh)
    float c2;       // 16 r^2: squared cut-off distance (the bump is dropped beyond 4 r)
    float h;        // 0.5 / r^2
    float a;        // signed amplitude [m/s]
};

// Geometry of the blob field of one padded grid.
struct RbShape {
    float x_max, z_max;   // (nxx - 1) dh, (nzz - 1) dh
    float pad;            // nb dh
    float ratio;          // minimum blob distance and largest radius [m]
    float amp;            // varVp [m/s]
    float bin;            // 4 ratio: side of the square bins of the table
    int nbx, nbz;         // bins along x and z (0 without blobs)
    int gx, gz;           // sampler grid cells along x and z
};

RbShape rb_shape(int nzz, int nxx, int nb, float dh, float ratio, float amp);

// Upper bound of the number of blobs of one stream (points at least `ratio`
// apart inside the outer half of the pad); 0 without blobs.
long rb_blob_bound(const RbShape &s);

// Buffers of the sampling on one host thread, reusable: reserved for a
// shape (reserve), sampling one stream with it allocates nothing.
struct RbWork {
    std::vector<int32_t> grid;
    std::vector<float> px, pz;
    std::vector<int32_t> active;
    std::vector<RbBlob> blobs;      // blobs of the streams sampled since clear()
    std::vector<long> counts;       // blobs per stream, in order
    void reserve(const RbShape &s);
    void clear() { blobs.clear(); counts.clear(); }
};

// Blobs of one stream in sample order, appended to w.blobs (and their count
// to w.counts).
void rb_blobs(const RbShape &s, uint64_t seed, uint64_t stream, RbWork &w);

// Blob table of streams[0 .. n): per stream, its blobs ordered by bin
// (b = bz + bx * nbz, bx = min((int)(x / bin), nbx - 1), bz likewise) and
// in sample order within a bin, written to blob[]; start[k * (nbins + 1) + b]
// is the absolute offset of bin b of stream k.  `work` holds one RbWork per
// host thread used (work.size() threads, at least one).  Returns the number
// of blobs written, or -1 if they would exceed `cap`.
long rb_blob_table(const RbShape &s, uint64_t seed, const uint64_t *streams, int n,
                   std::vector<RbWork> &work, RbBlob *blob, long cap, int32_t *start);

}  // namespace fwi
