// rbc.cpp -- see rbc.h.  Compiled with -ffp-contract=off.
#include "rbc.h"

#include <algorithm>
#include <cmath>
#include <thread>

namespace fwi {

namespace {

constexpr uint64_t kG = 0x9E3779B97F4A7C15ULL;
constexpr uint64_t kK = 0xD1B54A32D192ED03ULL;

uint64_t mix64(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ULL;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z;
}

struct Draw {
    uint64_t state;
    float u()
    {
        state += kG;
        return (float)(mix64(state) >> 40) * (1.0f / 16777216.0f);
    }
};

bool in_ring(const RbShape &s, float xh, float zh, float h, float x, float z)
{
    return !((x > h) && (x < xh) && (z > h) && (z < zh));
}

int bin_of(float v, float bin, int nb)
{
    const int b = (int)(v / bin);
    return b < nb - 1 ? b : nb - 1;
}

}  // namespace

RbShape rb_shape(int nzz, int nxx, int nb, float dh, float ratio, float amp)
{
    RbShape s{};
    s.x_max = (float)(nxx - 1) * dh;
    s.z_max = (float)(nzz - 1) * dh;
    s.pad = (float)nb * dh;
    s.ratio = ratio;
    s.amp = amp;
    if (!(ratio > 0.0f) || !(amp > 0.0f)) return s;
    s.bin = 4.0f * ratio;
    s.nbx = (int)(s.x_max / s.bin) + 1;
    s.nbz = (int)(s.z_max / s.bin) + 1;
    const float cs = ratio * 0.70710678f;
    s.gx = (int)(s.x_max / cs) + 1;
    s.gz = (int)(s.z_max / cs) + 1;
    return s;
}

long rb_blob_bound(const RbShape &s)
{
    if (s.nbx == 0) return 0;
    const double r = s.ratio;
    // disks of radius r/2 around the points are disjoint and lie in the ring
    // grown by r/2
    const double outer = ((double)s.x_max + r) * ((double)s.z_max + r);
    const double ix = (double)s.x_max - s.pad - r, iz = (double)s.z_max - s.pad - r;
    const double inner = (ix > 0 && iz > 0) ? ix * iz : 0.0;
    return (long)((outer - inner) / (M_PI * r * r / 4.0)) + 1;
}

void RbWork::reserve(const RbShape &s)
{
    const long n = rb_blob_bound(s);
    grid.reserve((size_t)s.gx * s.gz);
    px.reserve(n);
    pz.reserve(n);
    active.reserve(n);
    blobs.reserve(n);
    counts.reserve(1);
}

void rb_blobs(const RbShape &s, uint64_t seed, uint64_t stream, RbWork &w)
{
    const size_t first = w.blobs.size();
    if (s.nbx == 0) {
        w.counts.push_back(0);
        return;
    }
    const uint64_t key = mix64(seed ^ mix64(stream + kG));
    Draw g{mix64(key ^ kK)};
    const float r = s.ratio;
    const float r2 = r * r;
    const float four_r2 = 4.0f * r2;
    const float two_r = 2.0f * r;
    const float cs = r * 0.70710678f;
    const float h = 0.5f * s.pad;
    const float xh = s.x_max - h, zh = s.z_max - h;
    w.grid.assign((size_t)s.gx * s.gz, -1);
    w.px.clear();
    w.pz.clear();
    w.active.clear();
    auto valid = [&](float x, float z) {
        if (!(x >= 0.0f && x < s.x_max && z >= 0.0f && z < s.z_max)) return false;
        if (!in_ring(s, xh, zh, h, x, z)) return false;
        const int cx = (int)(x / cs), cz = (int)(z / cs);
        for (int ox = std::max(cx - 2, 0); ox <= std::min(cx + 2, s.gx - 1); ++ox)
            for (int oz = std::max(cz - 2, 0); oz <= std::min(cz + 2, s.gz - 1); ++oz) {
                const int k = w.grid[(size_t)oz + (size_t)ox * s.gz];
                if (k < 0) continue;
                const float dx = w.px[k] - x;
                const float dz = w.pz[k] - z;
                const float d2 = dx * dx + dz * dz;
                if (d2 < r2) return false;
            }
        return true;
    };
    auto insert = [&](float x, float z) {
        const int k = (int)w.px.size();
        w.px.push_back(x);
        w.pz.push_back(z);
        w.grid[(size_t)(int)(z / cs) + (size_t)(int)(x / cs) * s.gz] = k;
        w.active.push_back(k);
    };
    for (int t = 0; t < 10000; ++t) {
        const float x = g.u() * s.x_max;
        const float z = g.u() * s.z_max;
        if (in_ring(s, xh, zh, h, x, z)) {
            insert(x, z);
            break;
        }
    }
    while (!w.active.empty()) {
        const int na = (int)w.active.size();
        int a = (int)(g.u() * (float)na);
        if (a > na - 1) a = na - 1;
        const int k = w.active[a];
        const float x0 = w.px[k], z0 = w.pz[k];
        bool found = false;
        for (int t = 0; t < 30 && !found; ++t) {
            float dx, dz, d2;
            do {
                const float ux = 2.0f * g.u();
                dx = (ux - 1.0f) * two_r;
                const float uz = 2.0f * g.u();
                dz = (uz - 1.0f) * two_r;
                d2 = dx * dx + dz * dz;
            } while (!(d2 >= r2 && d2 < four_r2));
            const float x = x0 + dx, z = z0 + dz;
            if (valid(x, z)) {
                insert(x, z);
                found = true;
            }
        }
        if (!found) {
            w.active[a] = w.active.back();
            w.active.pop_back();
        }
    }
    for (size_t k = 0; k < w.px.size(); ++k) {
        const float u1 = g.u();
        const float t1 = 0.8f * u1;
        const float f1 = 0.2f + t1;
        const float rr = r * f1;
        const float u2 = g.u();
        const float t2 = 0.5f * u2;
        const float f2 = 0.5f + t2;
        const float amp = s.amp * f2;
        const float u3 = g.u();
        const float rr2 = rr * rr;
        RbBlob b;
        b.x = w.px[k];
        b.z = w.pz[k];
        b.c2 = 16.0f * rr2;
        b.h = 0.5f / rr2;
        b.a = (u3 < (1.0f / 3.0f)) ? amp : -amp;
        w.blobs.push_back(b);
    }
    w.counts.push_back((long)(w.blobs.size() - first));
}

long rb_blob_table(const RbShape &s, uint64_t seed, const uint64_t *streams, int n,
                   std::vector<RbWork> &work, RbBlob *blob, long cap, int32_t *start)
{
    const int nbins = s.nbx * s.nbz;
    const int nw = std::max(1, std::min((int)work.size(), n));
    for (auto &w : work) w.clear();
    // streams [lo(t), lo(t + 1)) on thread t, in order
    auto lo = [&](int t) { return (int)((long)n * t / nw); };
    auto run = [&](int t) {
        for (int k = lo(t); k < lo(t + 1); ++k) rb_blobs(s, seed, streams[k], work[t]);
    };
    if (nw == 1) {
        run(0);
    } else {
        std::vector<std::thread> pool;
        for (int t = 1; t < nw; ++t) pool.emplace_back(run, t);
        run(0);
        for (auto &th : pool) th.join();
    }
    long off = 0;
    for (int t = 0; t < nw; ++t) {
        const RbWork &w = work[t];
        long src = 0;
        for (int k = lo(t); k < lo(t + 1); ++k) {
            const long cnt = w.counts[k - lo(t)];
            int32_t *st = start + (long)k * (nbins + 1);
            if (off + cnt > cap || off + cnt > INT32_MAX) return -1;
            // counting sort by bin, stable
            std::fill(st, st + nbins + 1, 0);
            for (long m = 0; m < cnt; ++m) {
                const RbBlob &b = w.blobs[src + m];
                ++st[bin_of(b.z, s.bin, s.nbz) + bin_of(b.x, s.bin, s.nbx) * s.nbz + 1];
            }
            st[0] = (int32_t)off;
            for (int b = 0; b < nbins; ++b) st[b + 1] += st[b];
            // st[b] is now the first slot of bin b; place, then shift back
            for (long m = 0; m < cnt; ++m) {
                const RbBlob &b = w.blobs[src + m];
                const int bb = bin_of(b.z, s.bin, s.nbz) + bin_of(b.x, s.bin, s.nbx) * s.nbz;
                blob[st[bb]++] = b;
            }
            for (int b = nbins; b > 0; --b) st[b] = st[b - 1];
            st[0] = (int32_t)off;
            src += cnt;
            off += cnt;
        }
    }
    return off;
}

}  // namespace fwi
