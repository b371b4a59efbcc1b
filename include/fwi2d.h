/*
 * fwi2d.h -- C ABI of the MI355X-native acoustic modeling / RTM / FWI engine.
 *
 * One opaque context per survey (grid, model, wavelet, geometry, observed
 * data, accumulators) owning its device buffers and one HIP stream.  All
 * pointers in these signatures are HOST pointers owned by the caller and
 * only read or written during the call.  Every entry point returns 0 on
 * success or a positive FWI_ERR_* code; fwi_last_error() returns the message
 * of the last failure on the calling thread.  A context is used by one host
 * thread at a time.
 *
 * The reference (phbastosa/FWI2D) has no plugin or FFI layer: its operator
 * surface is the Modeling / Migration / Inversion class methods and their
 * CUDA kernels.  Each entry point below cites the methods it replaces.
 * Arrays follow the reference's file layouts: models and images [nx][nz]
 * with z fastest; seismograms [shot][nrec][nt] with t fastest.
 */
#ifndef FWI2D_H
#define FWI2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FWI_OK          0
#define FWI_ERR_ARG     1   /* invalid argument or geometry             */
#define FWI_ERR_STATE   2   /* call order: a required set_* is missing  */
#define FWI_ERR_HIP     3   /* HIP runtime failure (message has detail) */
#define FWI_ERR_NOMEM   4   /* device or host allocation failed         */
#define FWI_ERR_NUMERIC 5   /* CFL violated, non-positive model, ...    */
#define FWI_ERR_INTERNAL 6  /* an engine extent check failed (a bug)     */

#define FWI_BOUNDARY_ABC 0  /* Cerjan sponge on the abc_nb-padded grid  */
#define FWI_BOUNDARY_RBC 1  /* seeded random boundary on the rbc_nb grid */

/* Kinds of random boundary (fwi_config.rbc_kind), both seeded by rbc_seed. */
#define FWI_RBC_UNIFORM  0  /* uniform perturbation of amplitude rbc_varVp +
                               rbc_varVp_rel * edge speed, ramped 0 -> 1 */
#define FWI_RBC_GAUSSIAN 1  /* the reference's boundary: Gaussian noise of
                               standard deviation rbc_varVp ramped 0 -> 1, plus
                               Poisson-disk Gaussian blobs of radius up to
                               rbc_ratio and amplitude up to rbc_varVp in the
                               outer half of the pad
                               (src/modeling/modeling.cu:431-531) */

#define FWI_RESET_IMAGE    1
#define FWI_RESET_GRADIENT 2
#define FWI_RESET_ILLUM    4

typedef struct fwi_ctx fwi_ctx;

typedef struct fwi_config {
    int32_t  nz, nx;        /* z_samples, x_samples (model samples)            */
    float    dh;            /* model_spacing [m]                               */
    int32_t  nt;            /* time_samples                                    */
    float    dt;            /* time_spacing [s]                                */
    int32_t  abc_nb;        /* sponge pad per side, (int)(abc_length/dh)+1     */
    float    abc_factor;    /* Cerjan factor                                   */
    int32_t  rbc_nb;        /* random-boundary pad per side (cells)            */
    float    rbc_varVp;     /* random-boundary amplitude [m/s]                 */
    float    rbc_varVp_rel; /* extra amplitude as a fraction of the edge speed */
    uint64_t rbc_seed;      /* seed of the random boundary                     */
    int32_t  max_batch;     /* shots resident on the device per pass (>= 1)    */
    int32_t  rbc_kind;      /* FWI_RBC_UNIFORM (0) or FWI_RBC_GAUSSIAN         */
    float    rbc_ratio;     /* blob distance and largest radius [m] (GAUSSIAN;
                               0: no blobs); rbc_varVp_rel is not used there   */
} fwi_config;

/* Message of the last failure on this thread ("" if none). */
const char *fwi_last_error(void);
/* Library version string. */
const char *fwi_version(void);

/* Creates a context and its stream.  Replaces the set_parameters chains of
 * Modeling (src/modeling/modeling.cu:3-15), Migration
 * (src/migration/migration.cu:3-36) and Inversion (src/inversion/inversion.cu:3-53)
 * for everything that is not read from files. */
int fwi_create(const fwi_config *cfg, fwi_ctx **out);
int fwi_destroy(fwi_ctx *ctx);

/* Host helper: Ricker + half-derivative wavelet, FP64 DFT, rounded to FP32.
 * Replaces Modeling::set_wavelet (modeling.cu:32-90); returns tlag. */
int fwi_ricker_wavelet(int32_t nt, float dt, float fmax, float *w_out, int32_t *tlag_out);

/* Uploads the source wavelet w[nt]; T = nt + tlag steps per pass.
 * Replaces the cudaMemcpy of Modeling::set_wavelet (modeling.cu:82-84). */
int fwi_set_wavelet(fwi_ctx *ctx, const float *w, int32_t tlag);

/* Source / receiver coordinates in metres.  Grid indices follow
 * (int)((x + 0.5 dh)/dh) + nb per padded grid.  Replaces
 * Geometry::set_parameters (src/geometry/geometry.cpp:3-43) and
 * Modeling::initialization (modeling.cu:252-277). */
int fwi_set_geometry(fwi_ctx *ctx, int32_t nsrc, const float *xsrc, const float *zsrc,
                     int32_t nrec, const float *xrec, const float *zrec);

/* Interior velocity model vp[nx][nz] (m/s).  Padded on the device by edge
 * replication for both grids.  Replaces Modeling::set_properties
 * (modeling.cu:152-188) and the model uploads of Inversion::linesearch /
 * update_model (inversion.cu:310-321, 344-354). */
int fwi_set_model(fwi_ctx *ctx, const float *vp);

/* Observed (or migration input) data of shots [shot0, shot0+nshots):
 * data[nshots][nrec][nt], kept resident on the device.  Replaces
 * Migration::set_seismic_source (migration.cu:83-88) and
 * Inversion::set_obs_data (inversion.cu:168-172). */
int fwi_set_observed(fwi_ctx *ctx, int32_t shot0, int32_t nshots, const float *data);

/* Forward modeling of shots [shot0, shot0+nshots) on the chosen boundary;
 * if seis_out != NULL it receives [nshots][nrec][nt] (sample 0 = 0).
 * Replaces Modeling::forward_solver + compute_pressure + compute_seismogram
 * + get_seismogram (modeling.cu:279-353). */
int fwi_modeling(fwi_ctx *ctx, int32_t boundary, int32_t shot0, int32_t nshots,
                 float *seis_out);

/* RTM of shots [shot0, shot0+nshots) with the observed data as input:
 * random-boundary forward, time reversal and receiver back-propagation,
 * image += S*R and illum += S^2 accumulated on the device.  Replaces
 * Migration::forward_propagation / backward_propagation, inject_seismogram
 * and cross_correlation (migration.cu:49-81, 117-180). */
int fwi_migration(fwi_ctx *ctx, int32_t shot0, int32_t nshots);

/* FWI gradient of shots [shot0, shot0+nshots): sponge-boundary modeling,
 * residual obs - cal with its FP64 sum of squares added to *sum_res2,
 * random-boundary forward (stream (iteration, shot)), adjoint back-propagation
 * and grad += dt R v^2 lap(S), illum += S^2.  Replaces the shot loop body of
 * Inversion::compute_gradient (inversion.cu:119-134) with get_cal_data,
 * set_adjoint_source, update_RBC, forward/backward_propagation,
 * inject_adjoint and build_gradient (:174-248, 381-444). */
int fwi_gradient(fwi_ctx *ctx, int32_t shot0, int32_t nshots, uint32_t iteration,
                 double *sum_res2);

/* FP64 sum of squared residuals of sponge-boundary modeling against the
 * observed data, added to *sum_res2.  Replaces the shot loop of
 * Inversion::linesearch (inversion.cu:323-341). */
int fwi_misfit(fwi_ctx *ctx, int32_t shot0, int32_t nshots, double *sum_res2);

/* Zeroes accumulators (FWI_RESET_* bit mask).  The illumination is cumulative
 * over every migration / gradient call until reset. */
int fwi_reset(fwi_ctx *ctx, int32_t what);

/* RTM output: image_out[nx][nz] = -d2/dz2 (image / illum), zero on the first
 * and last rows; raw_image / raw_illum (cropped, may be NULL).  Replaces
 * Migration::export_seismic (migration.cu:90-115). */
int fwi_get_image(fwi_ctx *ctx, float *image_out, float *raw_image, float *raw_illum);

/* FWI search direction: g = -d2/dz2 (grad / illum) below zmask, divided by
 * max|g| (returned in *gmax).  raw_grad already carries the dt factor.
 * Replaces the post-processing of Inversion::compute_gradient
 * (inversion.cu:138-165). */
int fwi_get_gradient(fwi_ctx *ctx, float zmask, float *grad_out, float *raw_grad,
                     float *raw_illum, double *gmax);

/* Parabola vertex through (alpha_k, f_k), k = 0..2, in FP64.  Replaces the
 * step formula of Inversion::optimization (inversion.cu:304-307).  Returns
 * FWI_ERR_NUMERIC for a degenerate parabola. */
int fwi_parabola_step(const double *alpha, const double *f, double *step);

/* Step length for the update m - step * g with a guard the reference lacks
 * (it applies the vertex of inversion.cu:304-307 unconditionally at
 * inversion.cu:349, which gives NaN velocities when m turns non-positive).
 * "Admissible" means every m[k] - step g[k] >= m_min (m_min = 1/v^2 of the
 * fastest velocity the caller can propagate, or 0 for positivity only).
 * Rule 0: the parabola vertex, when the parabola through (alpha_k, f_k) is
 * convex, the vertex lies in (0, 4 alpha_2] and is admissible.  Rule 1:
 * otherwise the trial alpha_1 or alpha_2 with the lower misfit below f_0 that
 * is admissible (the other one next).  Rule 2: otherwise step = 0.  f_1 or
 * f_2 may be +inf for a trial that could not be modeled.  m and g hold n
 * values; *rule receives the rule taken. */
int fwi_guarded_step(const double *alpha, const double *f, const float *m, const float *g,
                     int64_t n, double m_min, double *step, int32_t *rule);

/* Shot-sum accumulators on the random-boundary grid as [nxx][nzz] (z
 * fastest, pads included): what = FWI_RESET_IMAGE, FWI_RESET_GRADIENT or
 * FWI_RESET_ILLUM (one of them).  For checkpoint / resume of an inversion,
 * whose illumination accumulates over every gradient pass
 * (inversion.cu:51,441).  No reference counterpart: the reference cannot
 * resume. */
int fwi_get_accumulator(fwi_ctx *ctx, int32_t what, float *out);
int fwi_set_accumulator(fwi_ctx *ctx, int32_t what, const float *in);

/* Test hook: the random-boundary velocity [nxx][nzz] (rbc grid) that a pass
 * of shot `shot` in FWI iteration `iteration` uses (migration: iteration 0). */
int fwi_rb_velocity(fwi_ctx *ctx, int32_t shot, uint32_t iteration, float *vp_padded_out);

/* Test hook, host only (no device needed): the blobs of the FWI_RBC_GAUSSIAN
 * boundary of one shot and iteration under cfg, in sample order, five floats
 * each (x, z [m] on the padded grid, 16 r^2, 0.5 / r^2, signed amplitude).
 * Writes at most cap blobs; *count gets the number of blobs.  Replaces the
 * host part of Modeling::set_random_boundary (src/modeling/modeling.cu:431-465)
 * and poissonDiskSampling (src/admin/admin.cpp:115-206). */
int fwi_rb_blobs(const fwi_config *cfg, int32_t shot, uint32_t iteration, float *out,
                 int64_t cap, int64_t *count);

/* fwi_rb_velocity without a device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  The value is staged in a static
 * device buffer of the library, one per GPU (8 MiB, held until the process
 * ends), so the random-boundary grid may hold at most 2,097,152 cells
 * (nzz * nxx); a larger grid fails with FWI_ERR_ARG.  Calls that overlap on
 * one GPU (from several threads or contexts) share the buffer and race. */
int fwi_rb_velocity_noalloc(fwi_ctx *ctx, int32_t shot, uint32_t iteration,
                            float *vp_padded_out);

/* The configuration the context was created with. */
int fwi_get_config(fwi_ctx *ctx, fwi_config *out);

/* Source and receiver counts of the current geometry (0 before
 * fwi_set_geometry). */
int fwi_geometry_size(fwi_ctx *ctx, int32_t *nsrc, int32_t *nrec);

/* Padded grid of a boundary kind: nzz, nxx, nb. */
int fwi_grid_shape(fwi_ctx *ctx, int32_t boundary, int32_t *nzz, int32_t *nxx, int32_t *nb);

/* Blocks until the context's stream is idle. */
int fwi_synchronize(fwi_ctx *ctx);

/* Launch tuning of the step kernels (kernel 0 random-boundary forward, 1
 * reverse, 2 sponge-boundary forward, -1 all):
 * time steps per launch (1-4: temporal blocking; a pass's remainder runs as
 * shorter launches), cells per lane along z (1, 2 or 4; the three- and
 * four-step reverse kernels take 1 or 2; 16 selects the kernels on
 * 64 x 64-cell tiles held in LDS, the shape for one or a few shots),
 * columns per wave (with the LDS tiles: the tile columns each wave
 * updates, 4, 8 or 16), and hipGraph capture of the time loops (1 on, 0
 * off).  0 / negative keeps the current value.  Every shape
 * gives bit-identical results. */
int fwi_set_tuning(fwi_ctx *ctx, int32_t kernel, int32_t steps, int32_t vz, int32_t xw,
                   int32_t use_graph);
int fwi_get_tuning(fwi_ctx *ctx, int32_t kernel, int32_t *steps, int32_t *vz, int32_t *xw,
                   int32_t *use_graph);

/* Autotuner: times every candidate launch shape of the forward and reverse
 * kernels on the random-boundary grid and of the sponge-boundary forward
 * kernel (receivers sampled) on its grid, for each grid that is configured,
 * with nslots resident shots (`reps` launches each), times the three
 * fastest again with 3 x reps launches, and keeps the fastest of those;
 * returns the time per step (us) of the random-boundary forward and of the
 * reverse kernel (0 without a random-boundary grid).  Until it runs, or
 * fwi_set_tuning sets them, every kernel uses four steps per launch, 81
 * columns per wave and 2 rows per lane.
 * Candidates: 1-4 steps per launch x 1, 2 or 4 rows per lane x chunk
 * widths 16-128 columns, 4 and 8 for small batches, and the widths that make
 * a launch 1-4 full rounds of the waves the device holds at once, and the
 * 64 x 64-cell LDS tiles with 1-4 steps and 4, 8 or 16 waves per tile.  No
 * reference counterpart (the reference launches one fixed shape). */
int fwi_autotune(fwi_ctx *ctx, int32_t nslots, int32_t reps, double *fwd_us, double *rev_us);

/* Measured device bandwidth (GB/s, read + write) of a plain float4 copy of
 * nbytes, `reps` times: the practical HBM ceiling the step kernels are
 * compared with. */
int fwi_stream_copy(int64_t nbytes, int32_t reps, double *gbs);

/* Per-pass timing with HIP events recorded on the context stream around each
 * time loop.  kind: 0 forward step, 1 reverse step.  Returns the summed GPU
 * milliseconds, the number of step launches and the cell-updates counted
 * (shots x nzz x nxx x steps, a reverse step counting 2 per cell). */
int fwi_profile_enable(fwi_ctx *ctx, int32_t enable);
int fwi_profile_read(fwi_ctx *ctx, int32_t kind, double *ms, int64_t *launches,
                     double *cell_updates);

#ifdef __cplusplus
}
#endif
#endif
