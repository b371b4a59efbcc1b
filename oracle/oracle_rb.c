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
