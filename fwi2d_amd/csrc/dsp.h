// dsp.h -- host-side numerics of the engine (see dsp.cpp).
#pragma once
#include <vector>

namespace fwi {

int ricker_half_derivative(int nt, double dt, double fmax, std::vector<double> &w);
void cerjan_table(int nb, double factor, std::vector<float> &b1d);
void second_derivative_z(const float *num, const float *den, int nz, int nx, double dh,
                         int row_lo, std::vector<double> &out);
bool parabola_vertex(const double *a, const double *f, double *step);

}  // namespace fwi
