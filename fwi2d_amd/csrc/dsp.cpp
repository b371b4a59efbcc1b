// dsp.cpp -- host-side numerics of the engine: the source wavelet (own FP64
// DFT; the reference uses FFTW, src/modeling/modeling.cu:50-80), the Cerjan
// table, image / gradient post-processing and the parabola vertex.  Plain C++.
#include "dsp.h"

#include <cmath>
#include <vector>

namespace fwi {

namespace {
constexpr double kPi = 3.14159265358979323846;
}

// Ricker 1e5 (1 - 2a) e^{-a}, a = pi^3 fc^2 (n dt - t0)^2, t0 = 2 sqrt(pi)/fmax,
// fc = fmax/(3 sqrt(pi)), then multiplied in frequency by (i 2 pi f)^{1/2}
// (principal root: sqrt(2 pi f) e^{i pi/4}, DC -> 0) and transformed back
// with 1/nt (modeling.cu:32-90).  Only bins 0..nt/2 enter the inverse,
// as in a real-to-complex / complex-to-real transform pair.
int ricker_half_derivative(int nt, double dt, double fmax, std::vector<double> &w)
{
    const double t0 = 2.0 * std::sqrt(kPi) / fmax;
    const double fc = fmax / (3.0 * std::sqrt(kPi));
    // tlag in single precision, as the reference computes it (modeling.cu:36-39)
    const float t0f = 2.0f * std::sqrt((float)kPi) / (float)fmax;
    const int tlag = (int)(t0f / (float)dt) + 1;
    std::vector<double> s(nt);
    for (int n = 0; n < nt; ++n) {
        const double td = n * dt - t0;
        const double a = kPi * kPi * kPi * fc * fc * td * td;
        s[n] = 1e5 * (1.0 - 2.0 * a) * std::exp(-a);
    }
    // twiddle table e^{-2 pi i m / nt}
    std::vector<double> cs(nt), sn(nt);
    for (int m = 0; m < nt; ++m) {
        const double ang = 2.0 * kPi * (double)m / (double)nt;
        cs[m] = std::cos(ang);
        sn[m] = std::sin(ang);
    }
    const int half = nt / 2;
    std::vector<double> re(half + 1), im(half + 1);
    for (int k = 0; k <= half; ++k) {
        double sr = 0.0, si = 0.0;
        long idx = 0;
        for (int n = 0; n < nt; ++n) {
            sr += s[n] * cs[idx];
            si -= s[n] * sn[idx];
            idx += k;
            if (idx >= nt) idx -= nt;
        }
        const double f = k / (nt * dt);
        const double mag = std::sqrt(2.0 * kPi * f);
        const double cr = mag * std::cos(kPi / 4.0), ci = mag * std::sin(kPi / 4.0);
        re[k] = sr * cr - si * ci;
        im[k] = sr * ci + si * cr;
    }
    w.assign(nt, 0.0);
    for (int n = 0; n < nt; ++n) {
        double acc = re[0];
        long idx = n;
        for (int k = 1; k <= half; ++k) {
            if ((nt % 2 == 0) && k == half) {
                acc += re[k] * ((n % 2) ? -1.0 : 1.0);
            } else {
                acc += 2.0 * (re[k] * cs[idx] - im[k] * sn[idx]);   // Re(X e^{+i theta})
            }
            idx += n;
            if (idx >= nt) idx -= nt;
        }
        w[n] = acc / nt;
    }
    return tlag;
}

// b1d[k] = exp(-(factor (nb - k))^2) (modeling.cu:125-128)
void cerjan_table(int nb, double factor, std::vector<float> &b1d)
{
    b1d.assign(nb > 0 ? nb : 1, 1.0f);
    for (int k = 0; k < nb; ++k) {
        const double a = factor * (double)(nb - k);
        b1d[k] = (float)std::exp(-(a * a));
    }
}

// q = num/den per cell; out = -(q[i-1] - 2 q[i] + q[i+1]) / dh^2 for rows
// row_lo < i < nz-1, else 0 (migration.cu:98-111, inversion.cu:144-158).
void second_derivative_z(const float *num, const float *den, int nz, int nx, double dh,
                         int row_lo, std::vector<double> &out)
{
    const double idh2 = 1.0 / (dh * dh);
    const long n = (long)nz * nx;
    std::vector<double> q(n);
    for (long k = 0; k < n; ++k) q[k] = (double)num[k] / (double)den[k];
    out.assign(n, 0.0);
    for (int j = 0; j < nx; ++j)
        for (int i = 0; i < nz; ++i) {
            const long k = i + (long)j * nz;
            if (i > row_lo && i < nz - 1)
                out[k] = -1.0 * (q[k - 1] - 2.0 * q[k] + q[k + 1]) * idh2;
        }
}

// Vertex of the parabola through (a_k, f_k) (inversion.cu:304-307).
bool parabola_vertex(const double *a, const double *f, double *step)
{
    const double num = (a[1] * a[1] - a[2] * a[2]) * f[0] + (a[2] * a[2] - a[0] * a[0]) * f[1]
                     + (a[0] * a[0] - a[1] * a[1]) * f[2];
    const double den = (a[1] - a[2]) * f[0] + (a[2] - a[0]) * f[1] + (a[0] - a[1]) * f[2];
    const double scale = std::fabs((a[1] - a[2]) * f[0]) + std::fabs((a[2] - a[0]) * f[1])
                       + std::fabs((a[0] - a[1]) * f[2]);
    if (!(std::fabs(den) > 1e-14 * scale) || !std::isfinite(num)) return false;
    *step = 0.5 * (num / den);
    return std::isfinite(*step);
}

}  // namespace fwi
