/*
 * oracle_rb.c -- CPU statement of the seeded random-boundary generator
 * (test infrastructure only; see fwi_oracle.h).  Must be compiled with
 * -ffp-contract=off: every float operation below is one correctly rounded
 * IEEE-754 operation, so the GPU generator can be checked bit for bit.
 *
 * The reference draws its pad velocity from a clock-seeded cuRAND normal per
 * cell (src/modeling/modeling.cu:487-531) plus Poisson-disk blobs
 * (modeling.cu:431-485, src/admin/admin.cpp:115-206) and is not reproducible.  This
 * generator keeps its structure -- clamp-to-edge base value, a linear ramp
 * from the interior edge (0) to the outer edge (1) applied once for the x pad
 * and once for the z pad (modeling.cu:497-520), pad cells only, clamped to a
 * velocity range (:586-587) -- with a counter-based draw:
 *
 *   key(seed, stream)       = mix64(seed ^ mix64(stream + G))
 *   h(seed, stream, cell,k) = mix64(key + (2*cell + k + 1) * G)
 *   u                       = ((int)(h >> 40) - 2^23) * 2^-23   in [-1, 1)
 *
 * with G = 0x9E3779B97F4A7C15 and mix64 the SplitMix64 finaliser
 * (Steele, Lea, Flood, OOPSLA 2014).  cell = i + j*nzz on the padded grid,
 * k = 0 for the x ramp and 1 for the z ramp.  Per cell in the pad:
 *
 *   A   = amp_abs + amp_rel * base
 *   val = base [+ (fx * A) * u0 if x pad] [+ (fz * A) * u1 if z pad]
 *   val = min(max(val, vlo), vhi)
 *   fx  = 1 - j/nb  (left)   or 1 - (nxx-1-j)/nb (right), same for fz.
 */
#include <stdint.h>
#include "fwi_oracle.h"

#define OR_RB_GOLDEN 0x9E3779B97F4A7C15ULL

uint64_t or_rb_mix64(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ULL;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z;
}

static uint64_t or_rb_key(uint64_t seed, uint64_t stream)
{
    return or_rb_mix64(seed ^ or_rb_mix64(stream + OR_RB_GOLDEN));
}

static float or_rb_u(uint64_t key, uint64_t cell, int k)
{
    uint64_t h = or_rb_mix64(key + (2 * cell + (uint64_t)k + 1) * OR_RB_GOLDEN);
    int32_t s = (int32_t)(h >> 40) - 8388608;      /* [-2^23, 2^23) */
    return (float)s * (1.0f / 8388608.0f);         /* exact */
}

float or_rb_uniform(uint64_t seed, uint64_t stream, uint64_t cell, int k)
{
    return or_rb_u(or_rb_key(seed, stream), cell, k);
}

void or_rb_velocity(const float *vpad, int nzz, int nxx, int nb,
                    float amp_abs, float amp_rel, float vlo, float vhi,
                    uint64_t seed, uint64_t stream, float *out)
{
    const uint64_t key = or_rb_key(seed, stream);
    const float fnb = (float)nb;
    for (int j = 0; j < nxx; j++) {
        for (int i = 0; i < nzz; i++) {
            const uint64_t cell = (uint64_t)i + (uint64_t)j * (uint64_t)nzz;
            const float base = vpad[cell];
            const int xpad = (j < nb) || (j >= nxx - nb);
            const int zpad = (i < nb) || (i >= nzz - nb);
            if (!xpad && !zpad) { out[cell] = base; continue; }
            const float ra = amp_rel * base;
            const float A = amp_abs + ra;
            float val = base;
            if (xpad) {
                const int dj = (j < nb) ? j : (nxx - 1 - j);
                const float q = (float)dj / fnb;
                const float f = 1.0f - q;
                const float fa = f * A;
                const float d = fa * or_rb_u(key, cell, 0);
                val = val + d;
            }
            if (zpad) {
                const int di = (i < nb) ? i : (nzz - 1 - i);
                const float q = (float)di / fnb;
                const float f = 1.0f - q;
                const float fa = f * A;
                const float d = fa * or_rb_u(key, cell, 1);
                val = val + d;
            }
            if (val < vlo) val = vlo;
            if (val > vhi) val = vhi;
            out[cell] = val;
        }
    }
}

/* ---- the reference-style boundary (FWI_RBC_GAUSSIAN) -------------------
 *
 * The reference (src/modeling/modeling.cu:431-531, src/admin/admin.cpp:115-206)
 * draws Gaussian noise of standard deviation varVp ramped across the pad and
 * adds Gaussian blobs at Poisson-disk points of the outer half of the pad.
 * The product's seeded statement of it (fwi2d_amd/csrc/rbc.h for the blob
 * sampler, rbg_kernel in propagate.hip for the field) is restated here step
 * by step so that both can be checked bit for bit:
 *
 *   normal(cell, k) = (2 S + 12 - 786432) 2^-17, S the sum of the twelve
 *       16-bit fields of mix64(key + (6 cell + 3k + m + 1) G), m = 0, 1, 2
 *       (Irwin-Hall: the sum of twelve uniforms minus 6);
 *   pad cell: val = base [+ (fx varVp) normal(cell, 0)] [+ (fz varVp) normal(cell, 1)],
 *       then the blobs of the 3 x 3 bins around the cell (x-major, z inner,
 *       table order within a bin) with d^2 < 16 r^2: val += a expneg(d^2 * (0.5 / r^2)),
 *       then val = min(max(val, vlo), vhi);
 *   expneg(q) = 2^-n P(y), n = (int)(q log2 e), y = q - n ln 2 by two fmaf
 *       with ln 2 = 0.693145752 + 1.42860682e-6, P the degree-9 Taylor
 *       polynomial of e^-y by Horner with fmaf.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_RB_K 0xD1B54A32D192ED03ULL

typedef struct { float x_max, z_max, pad, ratio, amp, bin; int nbx, nbz, gx, gz; } or_rb_shape;

static or_rb_shape or_rb_make_shape(int nzz, int nxx, int nb, float dh, float ratio, float amp)
{
    or_rb_shape s;
    memset(&s, 0, sizeof s);
    s.x_max = (float)(nxx - 1) * dh;
    s.z_max = (float)(nzz - 1) * dh;
    s.pad = (float)nb * dh;
    s.ratio = ratio;
    s.amp = amp;
    if (!(ratio > 0.0f) || !(amp > 0.0f)) return s;
    s.bin = 4.0f * ratio;
    s.nbx = (int)(s.x_max / s.bin) + 1;
    s.nbz = (int)(s.z_max / s.bin) + 1;
    {
        const float cs = ratio * 0.70710678f;
        s.gx = (int)(s.x_max / cs) + 1;
        s.gz = (int)(s.z_max / cs) + 1;
    }
    return s;
}

static float or_rb_draw(uint64_t *state)
{
    *state += OR_RB_GOLDEN;
    return (float)(or_rb_mix64(*state) >> 40) * (1.0f / 16777216.0f);
}

static int or_rb_outside_inner(float x, float z, float h, float xh, float zh)
{
    return !((x > h) && (x < xh) && (z > h) && (z < zh));
}

/* Blobs of one stream in sample order: out[5k .. 5k+4] = x, z, 16 r^2,
 * 0.5 / r^2, signed amplitude.  Returns the count (at most cap written). */
long or_rb_blobs(int nzz, int nxx, int nb, float dh, float ratio, float amp, uint64_t seed,
                 uint64_t stream, float *out, long cap)
{
    const or_rb_shape s = or_rb_make_shape(nzz, nxx, nb, dh, ratio, amp);
    if (s.nbx == 0) return 0;
    uint64_t state = or_rb_mix64(or_rb_key(seed, stream) ^ OR_RB_K);
    const float r = s.ratio, r2 = r * r, four_r2 = 4.0f * r2, two_r = 2.0f * r;
    const float cs = r * 0.70710678f;
    const float h = 0.5f * s.pad, xh = s.x_max - h, zh = s.z_max - h;
    const long ncell = (long)s.gx * s.gz;
    int *grid = malloc(sizeof(int) * ncell);
    long maxp = ncell + 1, np = 0, na = 0;
    float *px = malloc(sizeof(float) * maxp), *pz = malloc(sizeof(float) * maxp);
    long *active = malloc(sizeof(long) * maxp);
    for (long k = 0; k < ncell; k++) grid[k] = -1;
    for (int t = 0; t < 10000; t++) {
        const float x = or_rb_draw(&state) * s.x_max;
        const float z = or_rb_draw(&state) * s.z_max;
        if (or_rb_outside_inner(x, z, h, xh, zh)) {
            px[np] = x; pz[np] = z;
            grid[(long)(int)(z / cs) + (long)(int)(x / cs) * s.gz] = (int)np;
            active[na++] = np++;
            break;
        }
    }
    while (na > 0) {
        long a = (long)(int)(or_rb_draw(&state) * (float)na);
        if (a > na - 1) a = na - 1;
        const long k = active[a];
        const float x0 = px[k], z0 = pz[k];
        int found = 0;
        for (int t = 0; t < 30 && !found; t++) {
            float dx, dz, d2;
            for (;;) {
                const float ux = 2.0f * or_rb_draw(&state);
                dx = (ux - 1.0f) * two_r;
                const float uz = 2.0f * or_rb_draw(&state);
                dz = (uz - 1.0f) * two_r;
                d2 = dx * dx + dz * dz;
                if (d2 >= r2 && d2 < four_r2) break;
            }
            const float x = x0 + dx, z = z0 + dz;
            int ok = (x >= 0.0f && x < s.x_max && z >= 0.0f && z < s.z_max) &&
                     or_rb_outside_inner(x, z, h, xh, zh);
            if (ok) {
                const int cx = (int)(x / cs), cz = (int)(z / cs);
                for (int ox = cx - 2; ox <= cx + 2 && ok; ox++)
                    for (int oz = cz - 2; oz <= cz + 2 && ok; oz++) {
                        if (ox < 0 || oz < 0 || ox >= s.gx || oz >= s.gz) continue;
                        const int q = grid[(long)oz + (long)ox * s.gz];
                        if (q < 0) continue;
                        const float ex = px[q] - x, ez = pz[q] - z;
                        const float e2 = ex * ex + ez * ez;
                        if (e2 < r2) ok = 0;
                    }
            }
            if (ok) {
                px[np] = x; pz[np] = z;
                grid[(long)(int)(z / cs) + (long)(int)(x / cs) * s.gz] = (int)np;
                active[na++] = np++;
                found = 1;
            }
        }
        if (!found) active[a] = active[--na];
    }
    for (long k = 0; k < np; k++) {
        const float u1 = or_rb_draw(&state);
        const float t1 = 0.8f * u1, f1 = 0.2f + t1, rr = r * f1;
        const float u2 = or_rb_draw(&state);
        const float t2 = 0.5f * u2, f2 = 0.5f + t2, am = s.amp * f2;
        const float u3 = or_rb_draw(&state);
        const float rr2 = rr * rr;
        if (k < cap) {
            float *o = out + 5 * k;
            o[0] = px[k]; o[1] = pz[k]; o[2] = 16.0f * rr2; o[3] = 0.5f / rr2;
            o[4] = (u3 < (1.0f / 3.0f)) ? am : -am;
        }
    }
    free(grid); free(px); free(pz); free(active);
    return np;
}

float or_rb_normal(uint64_t seed, uint64_t stream, uint64_t cell, int k)
{
    const uint64_t key = or_rb_key(seed, stream);
    uint32_t sum = 0;
    for (int m = 0; m < 3; m++) {
        const uint64_t hh = or_rb_mix64(key + (6 * cell + 3 * (uint64_t)k + (uint64_t)m + 1) *
                                              OR_RB_GOLDEN);
        sum += (uint32_t)(hh & 0xFFFF) + (uint32_t)((hh >> 16) & 0xFFFF) +
               (uint32_t)((hh >> 32) & 0xFFFF) + (uint32_t)(hh >> 48);
    }
    return (float)((int32_t)(2 * sum) + 12 - 786432) * (1.0f / 131072.0f);
}

float or_rb_expneg(float q)
{
    const int n = (int)(q * 1.44269504f);
    const float fn = (float)n;
    float y = fmaf(-fn, 0.693145752f, q);      /* q - n ln 2, ln 2 split in two */
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

static int or_rb_bin(float v, float bin, int nb)
{
    const int b = (int)(v / bin);
    return b < nb - 1 ? b : nb - 1;
}

void or_rb_velocity_gauss(const float *vpad, int nzz, int nxx, int nb, float dh, float varVp,
                          float ratio, float vlo, float vhi, uint64_t seed, uint64_t stream,
                          float *out)
{
    const or_rb_shape s = or_rb_make_shape(nzz, nxx, nb, dh, ratio, varVp);
    const float fnb = (float)nb;
    long nbl = 0;
    float *bl = NULL;
    long *order = NULL, *start = NULL;
    const int nbins = s.nbx * s.nbz;
    if (s.nbx > 0) {
        nbl = or_rb_blobs(nzz, nxx, nb, dh, ratio, varVp, seed, stream, NULL, 0);
        bl = malloc(sizeof(float) * 5 * (nbl + 1));
        or_rb_blobs(nzz, nxx, nb, dh, ratio, varVp, seed, stream, bl, nbl);
        /* stable order by bin (b = bz + bx nbz) */
        start = calloc(nbins + 1, sizeof(long));
        order = malloc(sizeof(long) * (nbl + 1));
        for (long k = 0; k < nbl; k++)
            start[or_rb_bin(bl[5 * k + 1], s.bin, s.nbz) + or_rb_bin(bl[5 * k], s.bin, s.nbx) * s.nbz + 1]++;
        for (int b = 0; b < nbins; b++) start[b + 1] += start[b];
        {
            long *fill = malloc(sizeof(long) * (nbins + 1));
            memcpy(fill, start, sizeof(long) * (nbins + 1));
            for (long k = 0; k < nbl; k++)
                order[fill[or_rb_bin(bl[5 * k + 1], s.bin, s.nbz) + or_rb_bin(bl[5 * k], s.bin, s.nbx) * s.nbz]++] = k;
            free(fill);
        }
    }
    for (int j = 0; j < nxx; j++) {
        for (int i = 0; i < nzz; i++) {
            const uint64_t cell = (uint64_t)i + (uint64_t)j * (uint64_t)nzz;
            const float base = vpad[cell];
            const int xpad = (j < nb) || (j >= nxx - nb);
            const int zpad = (i < nb) || (i >= nzz - nb);
            if (!xpad && !zpad) { out[cell] = base; continue; }
            float val = base;
            if (xpad) {
                const int dj = (j < nb) ? j : (nxx - 1 - j);
                const float q = (float)dj / fnb;
                const float f = 1.0f - q;
                const float fs = f * varVp;
                const float d = fs * or_rb_normal(seed, stream, cell, 0);
                val = val + d;
            }
            if (zpad) {
                const int di = (i < nb) ? i : (nzz - 1 - i);
                const float q = (float)di / fnb;
                const float f = 1.0f - q;
                const float fs = f * varVp;
                const float d = fs * or_rb_normal(seed, stream, cell, 1);
                val = val + d;
            }
            if (nbins > 0) {
                const float x = (float)j * dh, z = (float)i * dh;
                const int cbx = or_rb_bin(x, s.bin, s.nbx), cbz = or_rb_bin(z, s.bin, s.nbz);
                for (int bx = cbx - 1; bx <= cbx + 1; bx++)
                    for (int bz = cbz - 1; bz <= cbz + 1; bz++) {
                        if (bx < 0 || bz < 0 || bx >= s.nbx || bz >= s.nbz) continue;
                        const int b = bz + bx * s.nbz;
                        for (long m = start[b]; m < start[b + 1]; m++) {
                            const float *B = bl + 5 * order[m];
                            const float dx = x - B[0], dz = z - B[1];
                            const float dx2 = dx * dx, dz2 = dz * dz;
                            const float d2 = dx2 + dz2;
                            if (d2 < B[2]) {
                                const float e = or_rb_expneg(d2 * B[3]);
                                const float add = B[4] * e;
                                val = val + add;
                            }
                        }
                    }
            }
            if (val < vlo) val = vlo;
            if (val > vhi) val = vhi;
            out[cell] = val;
        }
    }
    free(bl); free(order); free(start);
}
