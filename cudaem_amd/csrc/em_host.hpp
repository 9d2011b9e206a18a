// em_host.hpp -- host-side physics of the engine: constants, element tables,
// spline builder, slicing, lens derivation.  Double precision throughout; the
// device only sees the FP32 / FP64 results.
#pragma once
#include <cstdint>
#include <vector>

#include "cudaem_amd.h"
#include "em_kernels.hpp"

namespace em {

struct Constants {
  double lambda, gamma, sigma;
};
Constants beam_constants(double e0_kv);

// Kirkland-form projected-potential coefficients of one element.
struct ElementCoefs {
  std::vector<double> a, b, c, d;
};
ElementCoefs moliere_coefs(int Z);
void projected_potential(const ElementCoefs& k, double R, double* V, double* dVdR);

struct Spline {
  std::vector<double> R2, c0, c1, c2, c3;
  double rmax2;
};
Spline build_spline(const ElementCoefs& k, int nR, double r_min, double r_max, double tap_frac);

struct SliceRange {
  double z_start, z_end;
  long long i0, i1;  // inclusive atom range; i0 > i1 when empty
};
// z sorted ascending.  Slicing runs in float, as Slice_Generator<float> does:
// z, dz, eps and r_int are rounded to float and every bound and comparison is
// float; the bounds are returned as the float values, widened to double.
std::vector<SliceRange> make_slices(const std::vector<double>& z, double dz, int border, double eps);
// Fine-sliced mode: bounds over [z_min - r_int, z_max + r_int]; slice k holds
// the atoms within r_int of its centre (and those inside it), widened by eps.
std::vector<SliceRange> make_slices_fine(const std::vector<double>& z, double dz, int border, double eps,
                                         double r_int);

// Gauss-Legendre nodes and weights on [-1/2, 1/2], weights summing to 1.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);
// Gauss-Hermite nodes and weights for the weight exp(-t^2), ascending.
void gauss_hermite(int n, std::vector<double>& x, std::vector<double>& w);
// Moliere's screened potential as a 3D Kirkland 0-12 form (type 4):
// cl_i = e^2/(4 pi eps0) Z alpha_i, cnl_i = beta_i / a_TF, no Gaussians.
void moliere_3d(int Z, double cl[6], double cnl[6]);

LensDev derive_lens(const em_lens& l, double lambda, double defocus_shift);
// aperture / detector semi-angle (mrad) -> g^2 (1/A^2); outer <= 0 -> +inf
double mrad_to_g2(double mrad, double lambda, bool outer);

GridDev make_grid_dev(const em_grid& g);

}  // namespace em
