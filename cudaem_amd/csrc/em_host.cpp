// em_host.cpp -- host-side physics of the engine.
#include "em_host.hpp"

#include <algorithm>
#include <cmath>

namespace em {

namespace {
constexpr double kPi = 3.141592653589793238462643383279502884;
// CODATA 2018
constexpr double kHC = 12398.419843320026;   // eV * A
constexpr double kMC2 = 510998.95;           // eV
constexpr double kA0 = 0.529177210903;       // A
constexpr double kEk = 14.399645478425668;   // e^2 / (4 pi eps0), eV * A

// K0 and K1 from K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt by the
// trapezoid rule, exponentially convergent for this even analytic integrand.
void bessel_k01(double x, double* k0, double* k1) {
  const double h = 0.02;
  double s0 = 0.5 * std::exp(-x), s1 = s0;
  for (int k = 1; k < 200000; ++k) {
    const double t = k * h;
    const double c = std::cosh(t);
    if (x * c - t > 745.0) break;
    const double f = std::exp(-x * c);
    s0 += f;
    s1 += f * c;
  }
  *k0 = s0 * h;
  *k1 = s1 * h;
}
}  // namespace

// Relativistic constants (the reference reads them from the missing
// multislice_params.hpp, multislice_params_dev.cuh:131-133): textbook forms.
Constants beam_constants(double e0_kv) {
  const double e = e0_kv * 1e3;
  Constants c;
  c.lambda = kHC / std::sqrt(e * (2.0 * kMC2 + e));
  c.gamma = 1.0 + e / kMC2;
  c.sigma = 2.0 * kPi / (c.lambda * e) * (kMC2 + e) / (2.0 * kMC2 + e);
  return c;
}

// Moliere's screening of the Thomas-Fermi atom written in Kirkland's
// projected-potential form: a_i = Z alpha_i / (2 pi^2 a0),
// b_i = (beta_i / (2 pi a_TF))^2 with a_TF = 0.88534 a0 Z^(-1/3).
ElementCoefs moliere_coefs(int Z) {
  static const double alpha[3] = {0.35, 0.55, 0.10};
  static const double beta[3] = {0.3, 1.2, 6.0};
  const double atf = 0.88534 * kA0 * std::pow((double)Z, -1.0 / 3.0);
  ElementCoefs k;
  for (int i = 0; i < 3; ++i) {
    k.a.push_back(Z * alpha[i] / (2.0 * kPi * kPi * kA0));
    k.b.push_back(std::pow(beta[i] / (2.0 * kPi * atf), 2));
  }
  return k;
}

// VR(R) = 4 pi^2 a0 e sum a_i K0(2 pi R sqrt(b_i)) + 2 pi^2 a0 e sum c_i/d_i exp(-pi^2 R^2/d_i)
void projected_potential(const ElementCoefs& k, double R, double* V, double* dVdR) {
  const double fk = 4.0 * kPi * kPi * kA0 * kEk, fg = 2.0 * kPi * kPi * kA0 * kEk;
  double v = 0.0, dv = 0.0;
  for (size_t i = 0; i < k.a.size(); ++i) {
    const double q = 2.0 * kPi * std::sqrt(k.b[i]);
    double k0, k1;
    bessel_k01(q * R, &k0, &k1);
    v += fk * k.a[i] * k0;
    dv -= fk * k.a[i] * q * k1;
  }
  for (size_t i = 0; i < k.c.size(); ++i) {
    const double e = std::exp(-kPi * kPi * R * R / k.d[i]);
    v += fg * k.c[i] / k.d[i] * e;
    dv += fg * k.c[i] / k.d[i] * e * (-2.0 * kPi * kPi * R / k.d[i]);
  }
  *V = v;
  *dVdR = dv;
}

// Cubic spline of VR in R^2 (the reference's host builder is missing; its
// device side evaluates c0..c3 on R^2 nodes, basic_functions_dev.cuh:68-89,
// and tapers with apply_tapering_dev, basic_functions_dev.cuh:90-101).
Spline build_spline(const ElementCoefs& k, int nR, double r_min, double r_max, double tap_frac) {
  Spline s;
  s.R2.resize(nR);
  s.c0.resize(nR);
  s.c1.resize(nR);
  s.c2.assign(nR, 0.0);
  s.c3.assign(nR, 0.0);
  const double r_tap = tap_frac * r_max;
  const double r2_tap = r_tap * r_tap, r2_max = r_max * r_max;
  const double tap_cf = 0.5 * kPi / (r2_max - r2_tap);
  const double dl = std::log(r_max / r_min) / (nR - 1);
  for (int i = 0; i < nR; ++i) {
    const double r = (i == nR - 1) ? r_max : r_min * std::exp(i * dl);
    double v, dvdr;
    projected_potential(k, r, &v, &dvdr);
    double dv = dvdr / (2.0 * r);
    const double x = r * r;
    if (r2_tap < x) {
      const double ang = tap_cf * (x - r2_tap);
      const double tap = std::cos(ang), dtap = std::sin(ang);
      dv = dv * tap - tap_cf * v * dtap;
      v *= tap;
    }
    s.R2[i] = x;
    s.c0[i] = v;
    s.c1[i] = dv;
  }
  for (int i = 0; i < nR - 1; ++i) {
    const double idR = 1.0 / (s.R2[i + 1] - s.R2[i]);
    const double m = (s.c0[i + 1] - s.c0[i]) * idR;
    const double n = s.c1[i] + s.c1[i + 1];
    s.c2[i] = (3.0 * m - n - s.c1[i]) * idR;
    s.c3[i] = (n - 2.0 * m) * idR * idR;
  }
  s.rmax2 = s.R2[nR - 1];
  return s;
}

// Slice_Generator::gene_bounds / get_slicing / fd_by_z
// (slice_generator.hpp:81-168) in the real type T of the generator: the
// reference instantiates only Slice_Generator<float> (specimen.cuh:57 through
// the float simulators of simulator.cu:7-49), so the bounds, eps and every
// membership comparison are float there, and an atom within a float rounding
// of a bound lands in the slice the float arithmetic puts it in.
// Direct mode: slice k takes the atoms with z_start <= z < z_end.  Fine mode
// (:88-93, 108-128): bounds over [z_min - r_int, z_max + r_int] (the
// integration range Atom_Coord_Data reports, from a missing header: r_int =
// the largest spline radius), and slice k takes the atoms with z in
// [min(z_mid - r_int, z_start) - eps, max(z_mid + r_int, z_end) + eps]
// (Inc_Borders).
namespace {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
template <typename T>
std::vector<SliceRange> slice_atoms(const std::vector<T>& z, T dz, int border, T eps, bool fine, T r_int) {
  std::vector<SliceRange> out;
  const long long n = (long long)z.size();
  if (n == 0 || !(dz > 0)) return out;
  const T vmin = fine ? T(z.front() - r_int) : z.front();
  const T vmax = fine ? T(z.back() + r_int) : z.back();
  const T span = vmax - vmin;
  const int nv = static_cast<int>(std::floor(T(span + eps) / dz)) + 2;
  std::vector<T> b(nv);
  if (border == EM_BORDER_MIN) {
    const T zb = vmin - T(2) * eps;
    for (int i = 0; i < nv; ++i) b[i] = zb + T(i) * dz;
  } else if (border == EM_BORDER_MAX) {
    const T zt = vmax + T(2) * eps;
    for (int i = 0; i < nv; ++i) b[nv - 1 - i] = zt - T(i) * dz;
  } else {
    const T zb = T(0.5) * (T(vmin + vmax) - T(T(nv - 1) * dz));
    for (int i = 0; i < nv; ++i) b[i] = zb + T(i) * dz;
  }
  out.resize(nv - 1);
  for (int s = 0; s < nv - 1; ++s) {
    SliceRange& r = out[s];
    r.z_start = b[s];
    r.z_end = b[s + 1];
    T zs = b[s], ze = b[s + 1];
    if (fine) {
      const T zm = T(0.5) * (b[s] + b[s + 1]);
      zs = T(std::fmin(T(zm - r_int), b[s]) - eps);
      ze = T(std::fmax(T(zm + r_int), b[s + 1]) + eps);
    }
    if (zs > ze || z[n - 1] < zs || ze < z[0]) {
      r.i0 = 1;
      r.i1 = 0;
      continue;
    }
    r.i0 = (zs <= z[0]) ? 0 : (long long)(std::lower_bound(z.begin(), z.end(), zs) - z.begin());
    r.i1 = (z[n - 1] <= ze) ? n - 1 : (long long)(std::lower_bound(z.begin(), z.end(), ze) - z.begin()) - 1;
    if (r.i0 > r.i1 || z[r.i1] < zs || ze < z[r.i0]) {
      r.i0 = 1;
      r.i1 = 0;
    }
  }
  return out;
}

std::vector<float> to_float(const std::vector<double>& z) {
  std::vector<float> f(z.size());
  for (size_t i = 0; i < z.size(); ++i) f[i] = (float)z[i];
  return f;
}
}  // namespace

std::vector<SliceRange> make_slices(const std::vector<double>& z, double dz, int border, double eps) {
  return slice_atoms<float>(to_float(z), (float)dz, border, (float)eps, false, 0.f);
}

std::vector<SliceRange> make_slices_fine(const std::vector<double>& z, double dz, int border, double eps,
                                         double r_int) {
  return slice_atoms<float>(to_float(z), (float)dz, border, (float)eps, true, (float)r_int);
}

// Newton iteration on the Legendre polynomial P_n from the usual cosine
// guesses; nodes mapped to [-1/2, 1/2] with weights halved.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  auto legendre = [n](double t, double* pn, double* dpn) {
    double p0 = 1.0, p1 = t;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) {
      p1 = t;
      p0 = 1.0;
    }
    *pn = p1;
    *dpn = n * (t * p1 - p0) / (t * t - 1.0);
  };
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double t = std::cos(kPi * (i + 0.75) / (n + 0.5)), pn, dpn;
    for (int it = 0; it < 100; ++it) {
      legendre(t, &pn, &dpn);
      const double dt = pn / dpn;
      t -= dt;
      if (std::fabs(dt) < 1e-16) break;
    }
    legendre(t, &pn, &dpn);
    const double wi = 2.0 / ((1.0 - t * t) * dpn * dpn);
    x[i] = -0.5 * t;
    x[n - 1 - i] = 0.5 * t;
    w[i] = 0.5 * wi;
    w[n - 1 - i] = 0.5 * wi;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

// Newton iteration on the orthonormal Hermite recurrence
// p_{j+1} = t sqrt(2/(j+1)) p_j - sqrt(j/(j+1)) p_{j-1}, p_0 = pi^(-1/4),
// from the usual asymptotic starting points (largest node first); weight
// 2 / p_n'(t)^2 with p_n' = sqrt(2 n) p_{n-1}.
void gauss_hermite(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const int m = (n + 1) / 2;
  std::vector<double> r(m), rw(m);
  double t = 0.0;
  for (int i = 0; i < m; ++i) {
    if (i == 0) t = std::sqrt(2.0 * n + 1.0) - 1.85575 * std::pow(2.0 * n + 1.0, -0.16667);
    else if (i == 1) t -= 1.14 * std::pow((double)n, 0.426) / t;
    else if (i == 2) t = 1.86 * t - 0.86 * r[0];
    else if (i == 3) t = 1.91 * t - 0.91 * r[1];
    else t = 2.0 * t - r[i - 2];
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 0.7511255444649425, p0 = 0.0;
      for (int j = 0; j < n; ++j) {
        const double p2 = t * std::sqrt(2.0 / (j + 1)) * p1 - std::sqrt((double)j / (j + 1)) * p0;
        p0 = p1;
        p1 = p2;
      }
      dp = std::sqrt(2.0 * n) * p0;
      const double dt = p1 / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-15) break;
    }
    r[i] = t;
    rw[i] = 2.0 / (dp * dp);
  }
  for (int i = 0; i < m; ++i) {
    x[i] = -r[i];
    w[i] = rw[i];
    x[n - 1 - i] = r[i];
    w[n - 1 - i] = rw[i];
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

void moliere_3d(int Z, double cl[6], double cnl[6]) {
  static const double alpha[3] = {0.35, 0.55, 0.10};
  static const double beta[3] = {0.3, 1.2, 6.0};
  const double atf = 0.88534 * kA0 * std::pow((double)Z, -1.0 / 3.0);
  for (int i = 0; i < 3; ++i) {
    cl[i] = kEk * Z * alpha[i];
    cnl[i] = beta[i] / atf;
  }
  for (int i = 3; i < 6; ++i) {
    cl[i] = 0.0;
    cnl[i] = 1.0;
  }
}

double mrad_to_g2(double mrad, double lambda, bool outer) {
  if (outer && !(mrad > 0)) return HUGE_VAL;
  if (!outer && !(mrad > 0)) return 0.0;
  const double g = std::sin(mrad * 1e-3) / lambda;
  return g * g;
}

// Lens_Dev::assign + set_defocus (lens_dev.cuh:74-163) with the coefficient
// scalings of its comments (:40-57).
LensDev derive_lens(const em_lens& p, double lambda, double defocus_shift) {
  LensDev l;
  const double l1 = lambda, l2 = l1 * l1, l3 = l2 * l1, l4 = l3 * l1, l5 = l4 * l1;
  const double c10 = p.c_10 + defocus_shift;
  l.cc10 = (std::fabs(c10) <= 1e-5) ? 0.0 : -kPi * c10 * l1;
  l.cc12 = -kPi * p.c_12 * l1;
  l.cc21 = -2.0 * kPi * p.c_21 * l2 / 3.0;
  l.cc23 = -2.0 * kPi * p.c_23 * l2 / 3.0;
  l.cc30 = -kPi * p.c_30 * l3 / 2.0;
  l.cc32 = -kPi * p.c_32 * l3 / 2.0;
  l.cc34 = -kPi * p.c_34 * l3 / 2.0;
  l.cc41 = -2.0 * kPi * p.c_41 * l4 / 5.0;
  l.cc43 = -2.0 * kPi * p.c_43 * l4 / 5.0;
  l.cc45 = -2.0 * kPi * p.c_45 * l4 / 5.0;
  l.cc50 = -kPi * p.c_50 * l5 / 3.0;
  l.cc52 = -kPi * p.c_52 * l5 / 3.0;
  l.cc54 = -kPi * p.c_54 * l5 / 3.0;
  l.cc56 = -kPi * p.c_56 * l5 / 3.0;
  l.phi12 = p.phi_12; l.phi21 = p.phi_21; l.phi23 = p.phi_23; l.phi32 = p.phi_32;
  l.phi34 = p.phi_34; l.phi41 = p.phi_41; l.phi43 = p.phi_43; l.phi45 = p.phi_45;
  l.phi52 = p.phi_52; l.phi54 = p.phi_54; l.phi56 = p.phi_56;
  l.m = p.m;
  l.g2_min = mrad_to_g2(p.inner_aper_mrad, lambda, false);
  l.g2_max = mrad_to_g2(p.outer_aper_mrad, lambda, true);
  auto nz = [](double v) { return std::fabs(v) > 1e-5; };
  l.phi_required = (nz(l.m) || nz(l.cc12) || nz(l.cc21) || nz(l.cc23) || nz(l.cc32) || nz(l.cc34) ||
                    nz(l.cc41) || nz(l.cc43) || nz(l.cc45) || nz(l.cc52) || nz(l.cc54) || nz(l.cc56))
                       ? 1
                       : 0;
  return l;
}

// Grid_2d_Dev::set_input_data (grid_dev.cuh:38-56) in FP64.
GridDev make_grid_dev(const em_grid& g) {
  GridDev d;
  d.nx = g.nx;
  d.ny = g.ny;
  d.dx = (float)(g.lx / g.nx);
  d.dy = (float)(g.ly / g.ny);
  d.dgx = 1.0 / g.lx;
  d.dgy = 1.0 / g.ly;
  const double gmax = std::min((g.nx / 2) * d.dgx, (g.ny / 2) * d.dgy);
  const double gl = 2.0 * gmax / 3.0;
  d.gl2_max = gl * gl;
  const double dg0 = 0.25, fg0 = 1e-2;
  d.alpha = std::log(1.0 / fg0 - 1.0) / (std::pow(gl + dg0, 2) - d.gl2_max);
  return d;
}

}  // namespace em
