/*
 * oracle_misc.c -- closed-form and host-side pieces of the CPU oracle
 * (test infrastructure only; see fwi_oracle.h).  All arithmetic is FP64.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "fwi_oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Direct DFT, X_k = sum_n x_n exp(-2 pi i k n / N), k = 0..N/2.
 * The twiddle angle uses (k*n mod N) so it stays accurate for large N. */
void or_dft_r2c(int n, const double *x, double *re, double *im)
{
    for (int k = 0; k <= n / 2; k++) {
        double sr = 0.0, si = 0.0;
        for (int m = 0; m < n; m++) {
            long km = ((long)k * m) % n;
            double a = 2.0 * M_PI * (double)km / (double)n;
            sr += x[m] * cos(a);
            si -= x[m] * sin(a);
        }
        re[k] = sr;
        im[k] = si;
    }
}

/* Hermitian inverse (unnormalised), reading only k = 0..N/2 like a c2r
 * transform: x_m = X_0 + 2 sum_{0<k<N/2} Re(X_k e^{+2 pi i k m/N})
 *                 [+ Re(X_{N/2}) (-1)^m for even N]. */
void or_dft_c2r(int n, const double *re, const double *im, double *x)
{
    const int half = n / 2;
    for (int m = 0; m < n; m++) {
        double s = re[0];
        for (int k = 1; k <= half; k++) {
            if ((n % 2 == 0) && k == half) {
                s += re[k] * ((m % 2) ? -1.0 : 1.0);
                continue;
            }
            long km = ((long)k * m) % n;
            double a = 2.0 * M_PI * (double)km / (double)n;
            s += 2.0 * (re[k] * cos(a) - im[k] * sin(a));
        }
        x[m] = s;
    }
}

/* Modeling::set_wavelet, src/modeling/modeling.cu:32-90: Ricker wavelet
 * 1e5 (1 - 2a) e^{-a}, a = pi^3 fc^2 (n dt - t0)^2, filtered in frequency by
 * (i 2 pi f)^{1/2} (DC -> 0) and transformed back with 1/nt. */
int or_wavelet(int nt, double dt_in, double fmax_in, double *w)
{
    /* the reference parses dt and fmax with std::stof (modeling.cu:25-27) */
    const double dt = (double)(float)dt_in, fmax = (double)(float)fmax_in;
    const double t0 = 2.0 * sqrt(M_PI) / fmax;
    const double fc = fmax / (3.0 * sqrt(M_PI));
    /* tlag in single precision, as modeling.cu:36-39 computes it */
    const float t0f = 2.0f * sqrtf((float)M_PI) / (float)fmax;
    const int tlag = (int)(t0f / (float)dt) + 1;
    if (nt < 1) return tlag;
    double *s = (double *)malloc((size_t)nt * sizeof(double));
    double *re = (double *)malloc((size_t)(nt / 2 + 1) * sizeof(double));
    double *im = (double *)malloc((size_t)(nt / 2 + 1) * sizeof(double));
    for (int k = 0; k < nt; k++) {
        double td = k * dt - t0;
        double arg = M_PI * M_PI * M_PI * fc * fc * td * td;
        s[k] = 1e5 * (1.0 - 2.0 * arg) * exp(-arg);
    }
    or_dft_r2c(nt, s, re, im);
    const double df = 1.0 / (nt * dt);
    for (int k = 0; k <= nt / 2; k++) {
        double f = k * df;                 /* only non-negative bins are read */
        double mag = sqrt(2.0 * M_PI * f); /* |i 2 pi f|^{1/2}, phase +pi/4 */
        double cr = mag * cos(M_PI / 4.0), ci = mag * sin(M_PI / 4.0);
        double xr = re[k], xi = im[k];
        re[k] = xr * cr - xi * ci;
        im[k] = xr * ci + xi * cr;
    }
    or_dft_c2r(nt, re, im, s);
    for (int k = 0; k < nt; k++) w[k] = s[k] / nt;
    free(s); free(re); free(im);
    return tlag;
}

/* Modeling::set_abc_dampers, modeling.cu:125-128 */
void or_abc_b1d(int nb, double factor, double *b1d)
{
    for (int k = 0; k < nb; k++) {
        double a = factor * (double)(nb - k);
        b1d[k] = exp(-(a * a));
    }
}

/* get_boundary_damper, modeling.cu:355-402: 1 inside, b1d on the four edges,
 * b2d = b1d[a] + b1d[b] - 1 in the corners (b2d built at :234-244). */
double or_damper(const double *b1d, int i, int j, int nzz, int nxx, int nb)
{
    const int inz = (i >= nb) && (i < nzz - nb);
    const int inx = (j >= nb) && (j < nxx - nb);
    const int a = (i < nb) ? i : (nzz - 1 - i);
    const int b = (j < nb) ? j : (nxx - 1 - j);
    if (inz && inx) return 1.0;
    if (inx) return b1d[a];
    if (inz) return b1d[b];
    return b1d[a] + b1d[b] - 1.0;
}

/* Modeling::expand_boundary, modeling.cu:190-218 */
void or_expand(const float *in, int nz, int nx, int nb, float *out)
{
    const int nzz = nz + 2 * nb, nxx = nx + 2 * nb;
    for (int j = 0; j < nx; j++)
        for (int i = 0; i < nz; i++)
            out[(i + nb) + (long)(j + nb) * nzz] = in[i + (long)j * nz];
    for (int i = 0; i < nb; i++)
        for (int j = nb; j < nxx - nb; j++) {
            out[i + (long)j * nzz] = out[nb + (long)j * nzz];
            out[(nzz - i - 1) + (long)j * nzz] = out[(nzz - nb - 1) + (long)j * nzz];
        }
    for (int i = 0; i < nzz; i++)
        for (int j = 0; j < nb; j++) {
            out[i + (long)j * nzz] = out[i + (long)nb * nzz];
            out[i + (long)(nxx - j - 1) * nzz] = out[i + (long)(nxx - nb - 1) * nzz];
        }
}

/* Modeling::reduce_boundary, modeling.cu:220-230 */
void or_reduce(const float *in, int nz, int nx, int nb, float *out)
{
    const int nzz = nz + 2 * nb;
    for (int j = 0; j < nx; j++)
        for (int i = 0; i < nz; i++)
            out[i + (long)j * nz] = in[(i + nb) + (long)(j + nb) * nzz];
}

/* Modeling::initialization, modeling.cu:257-269 (float arithmetic) */
int or_grid_index(float x, float dh, int nb)
{
    return (int)((x + 0.5f * dh) / dh) + nb;
}

/* Migration::export_seismic, migration.cu:98-111 */
void or_image_post(const double *image, const double *illum, int nz, int nx,
                   double dh_in, double *out)
{
    const double dh = (double)(float)dh_in;   /* float model_spacing */
    const double idh2 = 1.0 / (dh * dh);
    const long n = (long)nz * nx;
    double *q = (double *)malloc((size_t)n * sizeof(double));
    for (long k = 0; k < n; k++) q[k] = image[k] / illum[k];
    for (long k = 0; k < n; k++) {
        int i = (int)(k % nz);
        out[k] = 0.0;
        if (i > 0 && i < nz - 1)
            out[k] = -1.0 * (q[k - 1] - 2.0 * q[k] + q[k + 1]) * idh2;
    }
    free(q);
}

/* Inversion::compute_gradient post-processing, inversion.cu:144-165 */
void or_gradient_post(const double *grad, const double *illum, int nz, int nx,
                      double dh_in, double zmask, double *out)
{
    const double dh = (double)(float)dh_in;   /* float model_spacing */
    const double idh2 = 1.0 / (dh * dh);
    const long n = (long)nz * nx;
    const int izm = (int)((float)zmask / (float)dh);   /* inversion.cu:157 in float */
    double *p = (double *)malloc((size_t)n * sizeof(double));
    for (long k = 0; k < n; k++) p[k] = grad[k] / illum[k];
    double gmax = 0.0;
    for (long k = 0; k < n; k++) {
        int i = (int)(k % nz);
        out[k] = 0.0;
        if (i > izm && i < nz - 1)
            out[k] = -1.0 * (p[k - 1] - 2.0 * p[k] + p[k + 1]) * idh2;
        if (fabs(out[k]) > gmax) gmax = fabs(out[k]);
    }
    for (long k = 0; k < n; k++) out[k] *= 1.0 / gmax;
    free(p);
}

/* Inversion::optimization, inversion.cu:304-307 */
double or_parabola_step(const double *a, const double *f)
{
    double num = (a[1] * a[1] - a[2] * a[2]) * f[0] + (a[2] * a[2] - a[0] * a[0]) * f[1]
               + (a[0] * a[0] - a[1] * a[1]) * f[2];
    double den = (a[1] - a[2]) * f[0] + (a[2] - a[0]) * f[1] + (a[0] - a[1]) * f[2];
    return 0.5 * (num / den);
}
