/*
 * fwi2d_update.h -- the model update of one FWI iteration, in C++ on the
 * host, over the engine (fwi2d.h) and the preconditioner network
 * (fwi2d_nn.h).
 *
 * fwi_update_model is the reference's Inversion::optimization, linesearch
 * and update_model (src/inversion/inversion.cu:286-358): two trial misfits
 * of the parabolic line search, the step (fwi_guarded_step) and
 * m <- m - step g.  Optionally it then tries a second search direction d
 * (no reference counterpart): the trial models m + k step d for the given
 * multiples k are forward-modeled, and the best one replaces the plain
 * update only when its misfit is below the forward-modeled misfit of the
 * plain update itself, so a trial can never leave a worse model than the
 * plain step.  All pointers are host pointers; models, gradients and
 * directions are [nx][nz] with z fastest.  Returns 0 or an FWI_ERR_* code
 * (fwi_last_error()).
 */
#ifndef FWI2D_UPDATE_H
#define FWI2D_UPDATE_H

#include <stdint.h>

#include "fwi2d.h"
#include "fwi2d_nn.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FWI_TRIAL_NONE    0  /* plain FWI: the reference's update, no extra pass      */
#define FWI_TRIAL_PLAIN   1  /* control: d = -g (the plain direction, longer steps)   */
#define FWI_TRIAL_NETWORK 2  /* d = fwi_trial_direction(net, ...)                      */
#define FWI_TRIAL_GIVEN   3  /* d supplied by the caller                               */

typedef struct fwi_update_opts {
    double alpha1, alpha2; /* line-search trial steps (0.02 and 0.05 of 1/vmin^2 - 1/vmax^2) */
    double m_min;          /* smallest admissible m (1/v^2 of the fastest stable velocity)    */
    float zmask;           /* depth_mask [m] of the trial direction (as the gradient's)       */
    int32_t trial;         /* FWI_TRIAL_*                                                      */
    fwi_nn *net;           /* FWI_TRIAL_NETWORK: batch >= 1, H = nx + (-nx mod 4),             */
                           /*                    W = nz + (-nz mod 4)                           */
    const float *direction;/* FWI_TRIAL_GIVEN: d [nx][nz]                                      */
    int32_t nmultiples;    /* 1 .. 4 trial multiples of the plain step                         */
    double multiples[4];
} fwi_update_opts;

typedef struct fwi_update_info {
    double f[3];           /* misfits sqrt(sum res^2) at alpha 0 (given), alpha1, alpha2       */
    double step;           /* plain step length                                                */
    int32_t rule;          /* rule of fwi_guarded_step (0 vertex, 1 best trial, 2 no step)     */
    double plain_misfit;   /* misfit of m - step g (modeled when a trial runs; else NaN)       */
    double trial_misfit;   /* best trial's misfit (+inf when no trial was modeled)             */
    double trial_multiple; /* its multiple k of the step (0 when none)                         */
    int32_t taken;         /* 0 plain update, 1 trial direction                                */
    int32_t passes;        /* misfit passes over all shots that this call ran                   */
} fwi_update_info;

/* One model update from the post-processed gradient g (fwi_get_gradient) at
 * the model m (slowness squared, updated in place) whose misfit is f0.
 * raw_grad and raw_illum (fwi_get_gradient's raw outputs) are read by
 * FWI_TRIAL_NETWORK only and may be NULL otherwise.  The misfit passes run
 * on shots [0, nsrc) with the context's current wavelet and observed data;
 * on return the context holds the updated model. */
int fwi_update_model(fwi_ctx *ctx, const fwi_update_opts *opts, double f0, const float *g,
                     const float *raw_grad, const float *raw_illum, float *m,
                     fwi_update_info *info);

/* The network's search direction: x = (raw_grad, raw_illum, m), each divided
 * by its maximum magnitude and edge-padded to the network's H x W; then
 * d = net(x) - g (cropped to nx x nz), divided by max |d|, and zero in the
 * rows i <= zmask / dh and i = nz - 1 (the gradient's depth mask). */
int fwi_trial_direction(fwi_nn *net, int32_t nx, int32_t nz, float dh, float zmask,
                        const float *raw_grad, const float *raw_illum, const float *m,
                        const float *g, float *d_out);

#ifdef __cplusplus
}
#endif
#endif
