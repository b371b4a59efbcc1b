// update.cpp -- the model update of one FWI iteration on the host
// (include/fwi2d_update.h): the reference's parabolic line search and model
// update (src/inversion/inversion.cu:286-358) over the engine's misfit
// passes, and the trial-direction rule of the learned preconditioner.
//
// Every model is slowness squared m = 1/v^2 in FP32, as the reference keeps
// it (inversion.cu:101-109); a trial model is m - a g (or m + a d) with the
// step a rounded to FP32 first, one multiply and one subtract per cell (no
// FMA: this file is compiled with -ffp-contract=off), and its velocity
// 1/sqrt(m) one correctly rounded square root and division per cell.
#include "fwi2d_update.h"
#include "status.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace {

using fwi::fail;

#define TRY(expr)                    \
    do {                             \
        int r_ = (expr);             \
        if (r_ != FWI_OK) return r_; \
    } while (0)

constexpr double kInf = std::numeric_limits<double>::infinity();

struct Survey {
    fwi_ctx *ctx;
    long n;              // cells of the interior model
    int32_t nsrc;
    float m_min;
    std::vector<float> v;   // velocity of the model being modeled
    int passes = 0;

    // sqrt(sum res^2) of model m over every shot; +inf when some cell is not
    // admissible (m <= 0, below m_min, or rejected by the engine's
    // stability check), which the step rule and the trial choice skip.
    int misfit(const float *m, double *out)
    {
        for (long k = 0; k < n; ++k) {
            if (!(m[k] > 0.f && m[k] >= m_min)) {
                *out = kInf;
                return FWI_OK;
            }
            v[k] = 1.0f / std::sqrt(m[k]);
        }
        const int rc = fwi_set_model(ctx, v.data());
        if (rc == FWI_ERR_NUMERIC) {
            *out = kInf;
            return FWI_OK;
        }
        TRY(rc);
        double s = 0.0;
        TRY(fwi_misfit(ctx, 0, nsrc, &s));
        ++passes;
        *out = std::sqrt(s);
        return FWI_OK;
    }

    // Leaves model m in the engine (the accepted model of the iteration).
    int install(const float *m)
    {
        for (long k = 0; k < n; ++k) v[k] = 1.0f / std::sqrt(m[k]);
        return fwi_set_model(ctx, v.data());
    }
};

// out = m + a d, with a rounded to FP32.
void along(const float *m, const float *d, double a, long n, float *out)
{
    const float af = (float)a;
    for (long k = 0; k < n; ++k) out[k] = m[k] + af * d[k];
}

// Zero rows i <= zmask / dh and the last row (the gradient's depth mask,
// inversion.cu:157).
void depth_mask(float *d, int nx, int nz, float dh, float zmask)
{
    const int izm = (int)(zmask / dh);
    for (int j = 0; j < nx; ++j) {
        float *col = d + (long)j * nz;
        for (int i = 0; i <= izm && i < nz; ++i) col[i] = 0.f;
        col[nz - 1] = 0.f;
    }
}

float max_abs(const float *a, long n)
{
    float s = 0.f;
    for (long k = 0; k < n; ++k) s = std::max(s, std::fabs(a[k]));
    return s;
}

}  // namespace

extern "C" {

int fwi_trial_direction(fwi_nn *net, int32_t nx, int32_t nz, float dh, float zmask,
                        const float *raw_grad, const float *raw_illum, const float *m,
                        const float *g, float *d_out)
{
    if (!net || !raw_grad || !raw_illum || !m || !g || !d_out)
        return fail(FWI_ERR_ARG, "null argument");
    if (nx < 1 || nz < 2 || !(dh > 0.f)) return fail(FWI_ERR_ARG, "bad grid %d x %d", nz, nx);
    int32_t nb = 0, H = 0, W = 0;
    TRY(fwi_nn_get_shape(net, &nb, &H, &W));
    const int px = (4 - nx % 4) % 4, pz = (4 - nz % 4) % 4;
    if (nb < 1 || H != nx + px || W != nz + pz)
        return fail(FWI_ERR_ARG, "network of %d x %d for a %d x %d model (needs %d x %d)", H, W,
                    nx, nz, nx + px, nz + pz);
    const long n = (long)nx * nz, npix = (long)H * W;
    // inputs: each channel over its maximum magnitude, edge-padded to H x W
    std::vector<float> x(3 * npix), y(npix);
    const float *chan[3] = {raw_grad, raw_illum, m};
    for (int c = 0; c < 3; ++c) {
        const float s = max_abs(chan[c], n);
        float *xc = x.data() + c * npix;
        for (int i = 0; i < H; ++i) {
            const float *src = chan[c] + (long)std::min(i, nx - 1) * nz;
            for (int j = 0; j < W; ++j) {
                const float a = src[std::min(j, nz - 1)];
                xc[(long)i * W + j] = s > 0.f ? a / s : a;
            }
        }
    }
    TRY(fwi_nn_forward(net, 1, x.data(), y.data()));
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nz; ++j) d_out[(long)i * nz + j] = y[(long)i * W + j] - g[(long)i * nz + j];
    const float s = max_abs(d_out, n);
    if (s > 0.f)
        for (long k = 0; k < n; ++k) d_out[k] = d_out[k] / s;
    depth_mask(d_out, nx, nz, dh, zmask);
    return FWI_OK;
}

int fwi_update_model(fwi_ctx *ctx, const fwi_update_opts *o, double f0, const float *g,
                     const float *raw_grad, const float *raw_illum, float *m,
                     fwi_update_info *info)
{
    if (!ctx || !o || !g || !m || !info) return fail(FWI_ERR_ARG, "null argument");
    if (o->trial < FWI_TRIAL_NONE || o->trial > FWI_TRIAL_GIVEN)
        return fail(FWI_ERR_ARG, "unknown trial kind %d", o->trial);
    if (o->trial != FWI_TRIAL_NONE && (o->nmultiples < 1 || o->nmultiples > 4))
        return fail(FWI_ERR_ARG, "1 to 4 trial multiples, not %d", o->nmultiples);
    if (!(o->alpha1 > 0.0 && o->alpha2 > o->alpha1) || !(o->m_min >= 0.0))
        return fail(FWI_ERR_ARG, "line-search steps must satisfy 0 < alpha1 < alpha2");
    if (o->trial == FWI_TRIAL_NETWORK && (!o->net || !raw_grad || !raw_illum))
        return fail(FWI_ERR_ARG, "a network trial needs the network and the raw gradient and "
                                 "illumination");
    if (o->trial == FWI_TRIAL_GIVEN && !o->direction)
        return fail(FWI_ERR_ARG, "a given trial direction is missing");
    if (!std::isfinite(f0) || f0 < 0.0) return fail(FWI_ERR_NUMERIC, "invalid misfit f0 %g", f0);
    fwi_config cfg{};
    TRY(fwi_get_config(ctx, &cfg));
    Survey sv{ctx, (long)cfg.nx * cfg.nz, 0, (float)o->m_min, {}};
    TRY(fwi_geometry_size(ctx, &sv.nsrc, nullptr));
    if (sv.nsrc < 1) return fail(FWI_ERR_STATE, "geometry not set (fwi_set_geometry)");
    const long n = sv.n;
    sv.v.resize(n);

    fwi_update_info r{};
    r.f[0] = f0;
    r.plain_misfit = std::numeric_limits<double>::quiet_NaN();
    r.trial_misfit = kInf;
    // line search: the trial models m - alpha g (inversion.cu:310-342)
    std::vector<float> plain(n), trial(n), best(n), d;
    const double alpha[3] = {0.0, o->alpha1, o->alpha2};
    for (int k = 1; k <= 2; ++k) {
        along(m, g, -alpha[k], n, trial.data());
        TRY(sv.misfit(trial.data(), &r.f[k]));
    }
    TRY(fwi_guarded_step(alpha, r.f, m, g, n, o->m_min, &r.step, &r.rule));
    along(m, g, -r.step, n, plain.data());          // the plain update (inversion.cu:349)
    const float *next = plain.data();

    if (o->trial != FWI_TRIAL_NONE && r.step > 0.0) {
        // the plain update's own misfit: a trial of the line search when the
        // step is one of its steps (the same model, bit for bit), else modeled
        if (r.step == alpha[1]) r.plain_misfit = r.f[1];
        else if (r.step == alpha[2]) r.plain_misfit = r.f[2];
        else TRY(sv.misfit(plain.data(), &r.plain_misfit));
        d.resize(n);
        if (o->trial == FWI_TRIAL_NETWORK) {
            TRY(fwi_trial_direction(o->net, cfg.nx, cfg.nz, cfg.dh, o->zmask, raw_grad, raw_illum,
                                    m, g, d.data()));
        } else {
            const float *src = o->trial == FWI_TRIAL_PLAIN ? g : o->direction;
            const float sign = o->trial == FWI_TRIAL_PLAIN ? -1.f : 1.f;
            for (long k = 0; k < n; ++k) d[k] = sign * src[k];
            depth_mask(d.data(), cfg.nx, cfg.nz, cfg.dh, o->zmask);
        }
        for (int q = 0; q < o->nmultiples; ++q) {
            along(m, d.data(), o->multiples[q] * r.step, n, trial.data());
            double f = kInf;
            TRY(sv.misfit(trial.data(), &f));
            if (f < r.trial_misfit) {
                r.trial_misfit = f;
                r.trial_multiple = o->multiples[q];
                best.swap(trial);
            }
        }
        if (std::isfinite(r.trial_misfit) && r.trial_misfit < r.plain_misfit) {
            next = best.data();
            r.taken = 1;
        }
    } else if (o->trial != FWI_TRIAL_NONE) {
        r.plain_misfit = f0;                          // no step: the model stays
    }
    std::copy(next, next + n, m);
    TRY(sv.install(m));
    r.passes = sv.passes;
    *info = r;
    return FWI_OK;
}

}  // extern "C"
