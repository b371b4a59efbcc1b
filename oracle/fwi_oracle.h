/*
 * fwi_oracle.h -- CPU restatement of phbastosa/FWI2D's modeling / RTM / FWI
 * hot path, used ONLY as test infrastructure: the parity tests, the smoke
 * check and the CPU-baseline leg of the benchmark call it as the checker.
 * No product code path links or calls anything declared here.
 *
 * Two precisions of the same code are exported:
 *   *_f64  the oracle proper (FP64 fields, FP64 arithmetic)
 *   *_f32  the same algorithm in FP32 (the CPU baseline of the stencil)
 *
 * Semantics follow the reference step by step; each function cites the
 * reference file:line it restates.  Where the reference is racy or
 * undefined, the oracle fixes one order and says which:
 *   - forward step: inject the source, stencil on the undamped current
 *     field, damp both fields, then sample receivers
 *     (src/modeling/modeling.cu:301-353);
 *   - seismogram sample 0 is 0 (modeling.cu:351 never writes it);
 *   - the random boundary is this project's seeded generator
 *     (oracle_rb.c), not the reference's clock-seeded cuRAND one.
 *
 * Layouts: every grid array is [x][z] with z fastest ("i + j*nzz"), the
 * seismogram / data arrays are [rec][t] with t fastest, as in the reference.
 */
#ifndef FWI_ORACLE_H
#define FWI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- closed-form pieces (oracle_misc.c) -------------------------------- */

/* Ricker + half-derivative source wavelet; modeling.cu:32-90.
 * w[nt] out (FP64), returns tlag = (int)(t0/dt)+1 with t0 = 2*sqrt(pi)/fmax. */
int or_wavelet(int nt, double dt, double fmax, double *w);

/* Direct real DFT pair used by or_wavelet (exposed for tests). */
void or_dft_r2c(int n, const double *x, double *re, double *im);    /* k = 0..n/2 */
void or_dft_c2r(int n, const double *re, const double *im, double *x); /* unnormalised */

/* Cerjan sponge table b1d[k] = exp(-(factor*(nb-k))^2); modeling.cu:113-150 */
void or_abc_b1d(int nb, double factor, double *b1d);
/* Region lookup of the sponge factor at padded cell (i,j); modeling.cu:355-402 */
double or_damper(const double *b1d, int i, int j, int nzz, int nxx, int nb);

/* Edge-replication padding and crop; modeling.cu:190-230 */
void or_expand(const float *in, int nz, int nx, int nb, float *out);
void or_reduce(const float *in, int nz, int nx, int nb, float *out);

/* Source / receiver index rule (int)((x+0.5dh)/dh)+nb; modeling.cu:252-277 */
int or_grid_index(float x, float dh, int nb);

/* -d2/dz2 of image/illum, zero on the top/bottom rows; migration.cu:90-115.
 * image, illum: cropped [nx][nz]; out [nx][nz] */
void or_image_post(const double *image, const double *illum, int nz, int nx,
                   double dh, double *out);
/* Gradient post-processing; inversion.cu:138-165.  grad, illum cropped. */
void or_gradient_post(const double *grad, const double *illum, int nz, int nx,
                      double dh, double zmask, double *out);
/* Parabola vertex through (a_k, f_k); inversion.cu:286-308 */
double or_parabola_step(const double *a, const double *f);

/* ---- seeded random boundary (oracle_rb.c) ------------------------------ */
uint64_t or_rb_mix64(uint64_t z);
float or_rb_uniform(uint64_t seed, uint64_t stream, uint64_t cell, int k);
/* vpad: edge-replicated padded model [nxx][nzz]; writes the RB model. */
void or_rb_velocity(const float *vpad, int nzz, int nxx, int nb,
                    float amp_abs, float amp_rel, float vlo, float vhi,
                    uint64_t seed, uint64_t stream, float *out);
/* The reference-style boundary (FWI_RBC_GAUSSIAN): blobs of one stream
 * (5 floats each, sample order; returns the count, writes at most cap),
 * the Irwin-Hall normal draw, e^-q, and the RB model. */
long or_rb_blobs(int nzz, int nxx, int nb, float dh, float ratio, float amp, uint64_t seed,
                 uint64_t stream, float *out, long cap);
float or_rb_normal(uint64_t seed, uint64_t stream, uint64_t cell, int k);
float or_rb_expneg(float q);
void or_rb_velocity_gauss(const float *vpad, int nzz, int nxx, int nb, float dh, float varVp,
                          float ratio, float vlo, float vhi, uint64_t seed, uint64_t stream,
                          float *out);

/* ---- propagators (oracle_core.inc, instantiated twice) ----------------- */
/* Laplacian (d2/dx2 + d2/dz2) of P on the updated cells, 0 on the 4-cell rim. */
void or_laplacian_f64(int nzz, int nxx, double dh, const double *P, double *out);
void or_laplacian_f32(int nzz, int nxx, double dh, const float *P, float *out);

/*
 * Forward pass (Modeling::forward_solver, modeling.cu:286-299, with
 * compute_pressure :405-449 and compute_seismogram :451-457).
 *   vp     padded velocity [nxx][nzz] (float, as the reference stores it)
 *   w      wavelet[nt] (already in the working precision)
 *   abc    1 = Cerjan sponge on (b1d[nb]), 0 = off
 *   seis   [nrec][nt] or NULL
 *   P_out  final current field u_T, Q_out final previous field (may be NULL)
 *   nthreads  OpenMP threads (0 = runtime default)
 */
void or_forward_f64(int nzz, int nxx, int nb, double dh, double dt, int nt, int tlag,
                    const float *vp, const double *w, int siz, int six,
                    int nrec, const int *riz, const int *rix,
                    int abc, const double *b1d,
                    double *seis, double *P_out, double *Q_out, int nthreads);
void or_forward_f32(int nzz, int nxx, int nb, double dh, double dt, int nt, int tlag,
                    const float *vp, const float *w, int siz, int six,
                    int nrec, const int *riz, const int *rix,
                    int abc, const double *b1d,
                    float *seis, float *P_out, float *Q_out, int nthreads);

/*
 * Reverse pass: source-wavefield reversal + receiver back-propagation
 * + imaging (mode 0, migration.cu:61-81,117-180) or gradient (mode 1,
 * inversion.cu:230-248,381-444).
 *   S_cur  = u_T   (reference d_P after the forward pass)
 *   S_old  = u_{T-1} (reference d_Pold)
 *   data   [nrec][nt] injected time-reversed at the receivers
 *   acc    image (mode 0) or gradient (mode 1), ACCUMULATED (+=), [nxx][nzz]
 *   illum  accumulated (+=) sum of S^2, [nxx][nzz]
 * S_cur/S_old are consumed (overwritten).
 */
void or_reverse_f64(int nzz, int nxx, double dh, double dt, int nt, int tlag,
                    const float *vp, double *S_cur, double *S_old,
                    int nrec, const int *riz, const int *rix, const double *data,
                    int mode, double *acc, double *illum, int nthreads);
void or_reverse_f32(int nzz, int nxx, double dh, double dt, int nt, int tlag,
                    const float *vp, float *S_cur, float *S_old,
                    int nrec, const int *riz, const int *rix, const float *data,
                    int mode, float *acc, float *illum, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
