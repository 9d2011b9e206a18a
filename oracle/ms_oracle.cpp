// ms_oracle.cpp -- FP64 host restatement of the cudaEM multislice hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: it is loaded
// by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
// by nothing else.  The product library (cudaem_amd/) never links, loads or
// calls it, and never falls back to it.
//
// orc_stem_f32 runs the STEM loop in FP32 as a second CPU timing; it is
// checked against the FP64 loop and is not an oracle.
//
// Every function restates the reference's algorithm in double precision
// (slicing, which is index work, in float as the reference does it) and
// cites the reference file:line it follows (the reference is a flat directory
// of CUDA headers).  Pieces the reference does not ship (element tables, the
// spline builder, sigma, the aperture conversion, the phonon RNG) are fixed
// here exactly as the product fixes them, and are flagged "parity unpinned
// against the reference" in the comment of each such function.
//
// Conventions (shared with the product):
//   * complex arrays are interleaved (re, im) doubles, row-major [row][col],
//     col (x) fastest;
//   * forward FFT = exp(-2 pi i nk/N), inverse = exp(+2 pi i nk/N), both
//     unnormalised (fft_dev.cuh:159-190);
//   * pixel (ix, iy) sits at (ix*dx, iy*dy); frequencies use the unshifted
//     index i < n/2 ? i : i - n (grid_dev.cuh:160-175).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

typedef std::complex<double> cd;
static const double kPi = 3.141592653589793238462643383279502884;

// CODATA 2018
static const double kHC = 12398.419843320026;   // h*c in eV*Angstrom
static const double kMC2 = 510998.95;           // electron rest energy, eV
static const double kA0 = 0.529177210903;       // Bohr radius, Angstrom
static const double kEk = 14.399645478425668;   // e^2/(4 pi eps0), eV*Angstrom

// ---------------------------------------------------------------------------
// Relativistic wavelength, gamma and interaction constant.  The reference reads
// them from the missing multislice_params.hpp (multislice_params_dev.cuh:131-133);
// these are the textbook closed forms (Kirkland, Advanced Computing in Electron
// Microscopy, eqs. 2.5 and 5.6).  Parity unpinned against the reference.
// ---------------------------------------------------------------------------
double wavelength(double e_kv) {
  const double e = e_kv * 1e3;
  return kHC / std::sqrt(e * (2.0 * kMC2 + e));
}
double gamma_factor(double e_kv) { return 1.0 + e_kv * 1e3 / kMC2; }
double sigma_factor(double e_kv) {
  const double e = e_kv * 1e3;
  return 2.0 * kPi / (wavelength(e_kv) * e) * (kMC2 + e) / (2.0 * kMC2 + e);
}

// ---------------------------------------------------------------------------
// Modified Bessel functions K0, K1 from the integral representation
// K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt, evaluated with the
// trapezoid rule (exponentially convergent for this analytic, even integrand).
// ---------------------------------------------------------------------------
void bessel_k01(double x, double* k0, double* k1) {
  const double h = 0.02;
  double s0 = 0.5 * std::exp(-x), s1 = 0.5 * std::exp(-x);
  for (int k = 1; k < 200000; ++k) {
    const double t = k * h;
    const double c = std::cosh(t);
    const double arg = x * c;
    if (arg - t > 745.0) break;
    const double f = std::exp(-arg);
    s0 += f;
    s1 += f * c;
  }
  *k0 = s0 * h;
  *k1 = s1 * h;
}

// ---------------------------------------------------------------------------
// Projected potential in Kirkland's functional form
//   VR(R) = 4 pi^2 a0 e sum_i a_i K0(2 pi R sqrt(b_i))
//         + 2 pi^2 a0 e sum_i (c_i/d_i) exp(-pi^2 R^2 / d_i)        [V*Angstrom]
// The reference's tables live in the missing element_coeffs_loader.hpp; the
// build ships Moliere's Thomas-Fermi screening (alpha = .35,.55,.10;
// beta = .3,1.2,6.0; a_TF = 0.88534 a0 Z^-1/3) written in this form:
//   a_i = Z alpha_i / (2 pi^2 a0),  b_i = (beta_i / (2 pi a_TF))^2.
// Parity unpinned against the reference's (missing) tables.
// ---------------------------------------------------------------------------
void moliere_coefs(int Z, double* a, double* b) {
  static const double alpha[3] = {0.35, 0.55, 0.10};
  static const double beta[3] = {0.3, 1.2, 6.0};
  const double atf = 0.88534 * kA0 * std::pow((double)Z, -1.0 / 3.0);
  for (int i = 0; i < 3; ++i) {
    a[i] = Z * alpha[i] / (2.0 * kPi * kPi * kA0);
    b[i] = std::pow(beta[i] / (2.0 * kPi * atf), 2);
  }
}

// The same Moliere screening as a 3D potential in the Kirkland 0-12 form
// (type 4): V(r) = sum_i e^2/(4 pi eps0) Z alpha_i exp(-beta_i r / a_TF) / r,
// i.e. cl_i = kEk Z alpha_i [V A], cnl_i = beta_i / a_TF [1/A], no Gaussians.
// Its projection along z is moliere_coefs' K0 sum: int V dz = 2 cl_i K0(cnl_i R).
void moliere_3d(int Z, double* cl, double* cnl) {
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

void vr_kirkland(int nk, const double* a, const double* b, int ng, const double* c,
                 const double* d, double R, double* V, double* dVdR) {
  const double fk = 4.0 * kPi * kPi * kA0 * kEk;
  const double fg = 2.0 * kPi * kPi * kA0 * kEk;
  double v = 0.0, dv = 0.0;
  for (int i = 0; i < nk; ++i) {
    const double q = 2.0 * kPi * std::sqrt(b[i]);
    double k0, k1;
    bessel_k01(q * R, &k0, &k1);
    v += fk * a[i] * k0;
    dv -= fk * a[i] * q * k1;
  }
  for (int i = 0; i < ng; ++i) {
    const double e = std::exp(-kPi * kPi * R * R / d[i]);
    v += fg * c[i] / d[i] * e;
    dv += fg * c[i] / d[i] * e * (-2.0 * kPi * kPi * R / d[i]);
  }
  *V = v;
  *dVdR = dv;
}

// ---------------------------------------------------------------------------
// Cubic spline of VR in R^2 on nR nodes, log-spaced in R from R_min to R_max.
// Node values c0 = VR, c1 = dVR/d(R^2); cosine taper in R^2 beyond R_tap
// exactly as apply_tapering_dev (basic_functions_dev.cuh:90-101, called at
// device_kernels.cuh:89-91); c2, c3 per interval as get_cubic_poly_coef
// (basic_functions_dev.cuh:76-89).  The node placement, R_min, R_max and R_tap
// are the build's choice (the reference's builder is missing): parity unpinned.
// ---------------------------------------------------------------------------
void build_spline(int nk, const double* a, const double* b, int ng, const double* c,
                  const double* d, double r_min, double r_max, double tap_frac, int nR,
                  double* R2, double* c0, double* c1, double* c2, double* c3) {
  const double r_tap = tap_frac * r_max;
  const double r2_tap = r_tap * r_tap, r2_max = r_max * r_max;
  const double tap_cf = 0.5 * kPi / (r2_max - r2_tap);
  const double dl = std::log(r_max / r_min) / (nR - 1);
  for (int i = 0; i < nR; ++i) {
    const double r = (i == nR - 1) ? r_max : r_min * std::exp(i * dl);
    double v, dvdr;
    vr_kirkland(nk, a, b, ng, c, d, r, &v, &dvdr);
    double dv = dvdr / (2.0 * r);  // dV/d(R^2)
    const double x = r * r;
    if (r2_tap < x) {  // apply_tapering_dev
      const double tap = std::cos(tap_cf * (x - r2_tap));
      const double dtap = std::sin(tap_cf * (x - r2_tap));
      dv = dv * tap - tap_cf * v * dtap;
      v *= tap;
    }
    R2[i] = x;
    c0[i] = v;
    c1[i] = dv;
  }
  for (int i = 0; i < nR - 1; ++i) {  // get_cubic_poly_coef
    const double idR = 1.0 / (R2[i + 1] - R2[i]);
    const double V = c0[i], Vn = c0[i + 1], dV = c1[i], dVn = c1[i + 1];
    const double m = (Vn - V) * idR;
    const double n = dV + dVn;
    c2[i] = (3.0 * m - n - dV) * idR;
    c3[i] = (n - 2.0 * m) * idR * idR;
  }
  c2[nR - 1] = 0.0;
  c3[nR - 1] = 0.0;
}

// Largest i in [0, nR-2] with R2[i] <= x (0 if x < R2[0]); the value the
// reference's unrolled 7-step search returns for nR = 128
// (basic_functions_dev.cuh:47-67).
static inline int spline_index(double x, const double* R2, int nR) {
  int i0 = 0, ie = nR - 1;
  while (ie - i0 > 1) {
    const int im = (i0 + ie) >> 1;
    if (x < R2[im]) ie = im; else i0 = im;
  }
  return i0;
}
static inline double spline_eval(double x, const double* R2, const double* c0,
                                 const double* c1, const double* c2, const double* c3,
                                 int nR) {
  // get_value_Cubic_Spline, basic_functions_dev.cuh:68-75
  const int ix = spline_index(x, R2, nR);
  const double dx = x - R2[ix];
  return ((c3[ix] * dx + c2[ix]) * dx + c1[ix]) * dx + c0[ix];
}

// ---------------------------------------------------------------------------
// Fine slicing (gen_sa_coeff_kernel + fill_spl_coeff_kernel,
// device_kernels.cuh:48-105, driven by projected_potential.cuh:73-127).
//
// 3D atomic potential V(r) and (1/r) dV/dr of the six parameterisations
// (Vr_dVrir_*, basic_functions_dev.cuh:104-187), coefficients cl (linear) and
// cnl (non-linear), both scaled by the quadrature weight f
// (Vr_dVrir_dev, basic_functions_dev.cuh:189-214):
//   1 Doyle-Turner 0-4, 2 Peng 0-4: 4 / 5 Gaussians cl e^{-cnl r^2}
//   3 Peng 0-12:                     5 Gaussians
//   4 Kirkland 0-12:                 3 Yukawa terms cl e^{-cnl r}/r + 3 Gaussians
//   5 Weickenmeier-Kohl 0-12:        6 terms cl erfc(cnl r)/r
//   6 Lobato 0-12:                   5 terms cl (2/(cnl r) + 1) e^{-cnl r}
// The type numbers are the build's (the reference's enum lives in a missing
// header); the coefficient tables are supplied by the caller.
// ---------------------------------------------------------------------------
void vr3d(int pt, double r, const double* cl, const double* cnl, double f, double* V, double* dVrir) {
  double v = 0.0, d = 0.0;
  const double r2 = r * r, ir = 1.0 / r;
  switch (pt) {
    case 1: case 2: case 3: {
      const int n = pt == 1 ? 4 : 5;
      for (int i = 0; i < n; ++i) {
        const double e = cl[i] * std::exp(-cnl[i] * r2);
        v += e;
        d += -2.0 * cnl[i] * e;
      }
      break;
    }
    case 4: {
      for (int i = 0; i < 3; ++i) {
        const double e = cl[i] * std::exp(-cnl[i] * r) * ir;
        v += e;
        d += -e * (cnl[i] + ir) * ir;
      }
      for (int i = 3; i < 6; ++i) {
        const double e = cl[i] * std::exp(-cnl[i] * r2);
        v += e;
        d += -2.0 * cnl[i] * e;
      }
      break;
    }
    case 5: {
      double g = 0.0;
      for (int i = 0; i < 6; ++i) {
        v += cl[i] * std::erfc(cnl[i] * r);
        g += cl[i] * cnl[i] * std::exp(-cnl[i] * cnl[i] * r2);
      }
      v *= ir;
      d = -(2.0 / std::sqrt(kPi) * g + v) / r2;
      break;
    }
    case 6: {
      double extra = 0.0;
      for (int i = 0; i < 5; ++i) {
        const double br = cnl[i] * r;
        const double e = cl[i] * std::exp(-br);
        v += e * (2.0 / br + 1.0);
        extra += e * (br + 1.0);
      }
      d = -(v + extra) / r2;
      break;
    }
    default:
      break;
  }
  *V = v * f;
  *dVrir = d * f;
}

// Gauss-Legendre nodes and weights on [-1/2, 1/2] (weights sum to 1), the
// quadrature gen_sa_coeff_kernel integrates over one slice with
// (z = dz x + a, weight dz w).  Newton iteration on P_n (textbook).
void gauss_legendre(int n, double* x, double* w) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double t = std::cos(kPi * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) { p1 = t; p0 = 1.0; }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      const double dt = p1 / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-16) break;
    }
    {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
    }
    const double wi = 2.0 / ((1.0 - t * t) * dp * dp);
    x[i] = -0.5 * t;
    x[n - 1 - i] = 0.5 * t;
    w[i] = 0.5 * wi;
    w[n - 1 - i] = 0.5 * wi;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

// Spline of the potential integrated over one slice of thickness dz whose
// centre sits at height a above the atom:
//   c0(R_i) = sum_q dz w_q V(sqrt((dz x_q + a)^2 + R_i^2)),
//   c1(R_i) = 1/2 sum_q dz w_q (1/r) dV/dr  (= d c0 / d(R^2))
// (gen_sa_coeff_kernel, device_kernels.cuh:48-94), then the R^2 taper of
// apply_tapering_dev and the cubic coefficients of get_cubic_poly_coef
// (fill_spl_coeff_kernel, device_kernels.cuh:95-105).  Nodes as build_spline.
void fine_spline(int pt, const double* cl, const double* cnl, double a, double dz, int nq,
                 double r_min, double r_max, double tap_frac, int nR, double* R2, double* c0,
                 double* c1, double* c2, double* c3) {
  std::vector<double> qx(nq), qw(nq);
  gauss_legendre(nq, qx.data(), qw.data());
  const double r_tap = tap_frac * r_max;
  const double r2_tap = r_tap * r_tap, r2_max = r_max * r_max;
  const double tap_cf = 0.5 * kPi / (r2_max - r2_tap);
  const double dl = std::log(r_max / r_min) / (nR - 1);
  for (int i = 0; i < nR; ++i) {
    const double r = (i == nR - 1) ? r_max : r_min * std::exp(i * dl);
    const double x = r * r;
    double v = 0.0, dv = 0.0;
    for (int q = 0; q < nq; ++q) {
      const double z = dz * qx[q] + a;
      double vq, dq;
      vr3d(pt, std::sqrt(z * z + x), cl, cnl, dz * qw[q], &vq, &dq);
      v += vq;
      dv += dq;
    }
    dv *= 0.5;
    if (r2_tap < x) {
      const double tap = std::cos(tap_cf * (x - r2_tap));
      const double dtap = std::sin(tap_cf * (x - r2_tap));
      dv = dv * tap - tap_cf * v * dtap;
      v *= tap;
    }
    R2[i] = x;
    c0[i] = v;
    c1[i] = dv;
  }
  for (int i = 0; i < nR - 1; ++i) {
    const double idR = 1.0 / (R2[i + 1] - R2[i]);
    const double m = (c0[i + 1] - c0[i]) * idR;
    const double n = c1[i] + c1[i + 1];
    c2[i] = (3.0 * m - n - c1[i]) * idR;
    c3[i] = (n - 2.0 * m) * idR * idR;
  }
  c2[nR - 1] = 0.0;
  c3[nR - 1] = 0.0;
}

// ---------------------------------------------------------------------------
// Mixed-radix FFT, unnormalised, sign -1 forward / +1 inverse.
// Recursive decimation in time over the prime factors of n (4 preferred).
// ---------------------------------------------------------------------------
// Mixed-radix recursive FFT in the real type R (double: the oracle; float:
// the FP32 variant timed beside it).  Twiddles are formed in double.
template <class R>
struct FftT {
  typedef std::complex<R> C;
  int n;
  std::vector<C> w;  // w[j] = exp(-2 pi i j / n)
  explicit FftT(int n_) : n(n_), w(n_) {
    for (int j = 0; j < n; ++j) {
      const double ang = -2.0 * kPi * (double)j / (double)n;
      w[j] = C((R)std::cos(ang), (R)std::sin(ang));
    }
  }
  static int first_factor(int m) {
    if (m % 4 == 0) return 4;
    if (m % 2 == 0) return 2;
    for (int p = 3; p * p <= m; p += 2)
      if (m % p == 0) return p;
    return m;
  }
  // out[k], k < m, = sum_j in[j*stride] exp(sign 2 pi i jk/m); m divides n
  void rec(const C* in, int stride, C* out, int m, int sign, C* tmp) const {
    if (m == 1) { out[0] = in[0]; return; }
    const int p = first_factor(m);
    const int q = m / p;
    for (int r = 0; r < p; ++r) rec(in + r * stride, stride * p, out + r * q, q, sign, tmp);
    const int step = n / m;  // W_m^x = w[x * step]
    for (int k = 0; k < q; ++k) {
      for (int r = 0; r < p; ++r) {
        const int e = (int)(((long long)r * k) % m) * step;
        const C tw = sign < 0 ? w[e] : std::conj(w[e]);
        tmp[r] = out[r * q + k] * tw;
      }
      if (p == 2) {
        out[k] = tmp[0] + tmp[1];
        out[k + q] = tmp[0] - tmp[1];
        continue;
      }
      if (p == 4) {
        // W_4 = exp(sign i pi / 2) = sign * i
        const C t0 = tmp[0] + tmp[2], t1 = tmp[0] - tmp[2], t2 = tmp[1] + tmp[3];
        const C d = tmp[1] - tmp[3];
        const C t3 = sign < 0 ? C(d.imag(), -d.real()) : C(-d.imag(), d.real());
        out[k] = t0 + t2;
        out[k + q] = t1 + t3;
        out[k + 2 * q] = t0 - t2;
        out[k + 3 * q] = t1 - t3;
        continue;
      }
      for (int s = 0; s < p; ++s) {
        C acc(0, 0);
        for (int r = 0; r < p; ++r) {
          const int e = (int)(((long long)r * s * q) % m) * step;
          acc += tmp[r] * (sign < 0 ? w[e] : std::conj(w[e]));
        }
        out[k + s * q] = acc;
      }
    }
  }
  void run(C* data, int stride, int sign, C* buf_in, C* buf_out, C* tmp) const {
    for (int j = 0; j < n; ++j) buf_in[j] = data[(long long)j * stride];
    rec(buf_in, 1, buf_out, n, sign, tmp);
    for (int j = 0; j < n; ++j) data[(long long)j * stride] = buf_out[j];
  }
};
typedef FftT<double> Fft;

// In-place 2D FFT of a [ny][nx] array: rows, then columns.
template <class R>
void fft2(std::complex<R>* a, int nx, int ny, int sign) {
  FftT<R> fx(nx), fy(ny);
  const int mx = std::max(nx, ny);
  std::vector<std::complex<R>> bi(mx), bo(mx), tmp(mx);
  for (int y = 0; y < ny; ++y) fx.run(a + (long long)y * nx, 1, sign, bi.data(), bo.data(), tmp.data());
  for (int x = 0; x < nx; ++x) fy.run(a + x, nx, sign, bi.data(), bo.data(), tmp.data());
}

// ---------------------------------------------------------------------------
// Grid helpers, grid_dev.cuh:38-56, 84-95, 160-196.
// ---------------------------------------------------------------------------
struct Grid {
  int nx, ny;
  double lx, ly, dz;
  int bwl;
  double dgx, dgy, dx, dy, gl2_max, alpha;
  Grid(int nx_, int ny_, double lx_, double ly_, double dz_, int bwl_)
      : nx(nx_), ny(ny_), lx(lx_), ly(ly_), dz(dz_), bwl(bwl_) {
    dx = lx / nx;
    dy = ly / ny;
    dgx = 1.0 / lx;
    dgy = 1.0 / ly;
    const double gmax = std::min((nx / 2) * dgx, (ny / 2) * dgy);
    const double gl = 2.0 * gmax / 3.0;
    gl2_max = gl * gl;
    const double dg0 = 0.25, fg0 = 1e-2;
    alpha = std::log(1.0 / fg0 - 1.0) / (std::pow(gl + dg0, 2) - gl2_max);
  }
  double gx(int ix) const { return (ix < nx / 2 ? ix : ix - nx) * dgx; }
  double gy(int iy) const { return (iy < ny / 2 ? iy : iy - ny) * dgy; }
  double bwl_factor(int ix, int iy) const {
    const double g2 = gx(ix) * gx(ix) + gy(iy) * gy(iy);
    return 1.0 / (1.0 + std::exp(alpha * (g2 - gl2_max)));
  }
  long long n() const { return (long long)nx * ny; }
};

// ---------------------------------------------------------------------------
// Lens: coefficient layout shared with the product (em_lens in
// include/cudaem_amd.h), lens_dev.cuh:12-66 with the scalings of its comments
// :40-57.  Apertures: semi-angle (mrad) -> g = sin(angle)/lambda; an outer
// aperture <= 0 means "no aperture" (the reference's export_g2_* are missing:
// parity unpinned).
// ---------------------------------------------------------------------------
enum {
  L_C10, L_C12, L_PHI12, L_C21, L_PHI21, L_C23, L_PHI23, L_C30, L_C32, L_PHI32, L_C34,
  L_PHI34, L_C41, L_PHI41, L_C43, L_PHI43, L_C45, L_PHI45, L_C50, L_C52, L_PHI52, L_C54,
  L_PHI54, L_C56, L_PHI56, L_M, L_INNER, L_OUTER, L_COUNT
};

struct Lens {
  double m, cc10, cc12, cc21, cc23, cc30, cc32, cc34, cc41, cc43, cc45, cc50, cc52, cc54, cc56;
  double phi12, phi21, phi23, phi32, phi34, phi41, phi43, phi45, phi52, phi54, phi56;
  double g2_min, g2_max;
  // isZero_dev<float> gate (isEqual_dev<float>, misc_dev.cuh:14-24, 36-53; used via nonZero_dev in lens_dev.cuh:165-252)
  static bool nz(double v) { return std::fabs(v) > 1e-5; }
  Lens(const double* p, double lambda, double defocus_shift) {
    const double l = lambda, l2 = l * l, l3 = l2 * l, l4 = l3 * l, l5 = l4 * l;
    const double c10 = p[L_C10] + defocus_shift;
    // set_defocus, lens_dev.cuh:159-163
    cc10 = (std::fabs(c10) <= 1e-5) ? 0.0 : -kPi * c10 * l;
    cc12 = -kPi * p[L_C12] * l;
    cc21 = -2.0 * kPi * p[L_C21] * l2 / 3.0;
    cc23 = -2.0 * kPi * p[L_C23] * l2 / 3.0;
    cc30 = -kPi * p[L_C30] * l3 / 2.0;
    cc32 = -kPi * p[L_C32] * l3 / 2.0;
    cc34 = -kPi * p[L_C34] * l3 / 2.0;
    cc41 = -2.0 * kPi * p[L_C41] * l4 / 5.0;
    cc43 = -2.0 * kPi * p[L_C43] * l4 / 5.0;
    cc45 = -2.0 * kPi * p[L_C45] * l4 / 5.0;
    cc50 = -kPi * p[L_C50] * l5 / 3.0;
    cc52 = -kPi * p[L_C52] * l5 / 3.0;
    cc54 = -kPi * p[L_C54] * l5 / 3.0;
    cc56 = -kPi * p[L_C56] * l5 / 3.0;
    phi12 = p[L_PHI12]; phi21 = p[L_PHI21]; phi23 = p[L_PHI23]; phi32 = p[L_PHI32];
    phi34 = p[L_PHI34]; phi41 = p[L_PHI41]; phi43 = p[L_PHI43]; phi45 = p[L_PHI45];
    phi52 = p[L_PHI52]; phi54 = p[L_PHI54]; phi56 = p[L_PHI56];
    m = p[L_M];
    const double gi = std::sin(p[L_INNER] * 1e-3) / lambda;
    g2_min = p[L_INNER] > 0 ? gi * gi : 0.0;
    const double go = std::sin(p[L_OUTER] * 1e-3) / lambda;
    g2_max = p[L_OUTER] > 0 ? go * go : HUGE_VAL;
  }
  bool phi_required() const {  // lens_dev.cuh:165-170
    return nz(m) || nz(cc12) || nz(cc21) || nz(cc23) || nz(cc32) || nz(cc34) || nz(cc41) ||
           nz(cc43) || nz(cc45) || nz(cc52) || nz(cc54) || nz(cc56);
  }
  // coherent_phase_shift, basic_functions_dev.cuh:215-234
  double chi(double gx, double gy) const {
    const double g2 = gx * gx + gy * gy, g4 = g2 * g2, g6 = g4 * g2;
    double ph = (nz(cc10) ? cc10 * g2 : 0.0) + (nz(cc30) ? cc30 * g4 : 0.0) +
                (nz(cc50) ? cc50 * g6 : 0.0);
    if (phi_required()) {
      const double g = std::sqrt(g2), g3 = g2 * g, g5 = g4 * g;
      const double phi = std::atan2(gy, gx);
      ph += nz(phi) ? m * phi : 0.0;
      ph += nz(cc12) ? cc12 * g2 * std::sin(2 * (phi - phi12)) : 0.0;
      ph += nz(cc21) ? cc21 * g3 * std::sin(phi - phi21) : 0.0;
      ph += nz(cc23) ? cc23 * g3 * std::sin(3 * (phi - phi23)) : 0.0;
      ph += nz(cc32) ? cc32 * g4 * std::sin(2 * (phi - phi32)) : 0.0;
      ph += nz(cc34) ? cc34 * g4 * std::sin(4 * (phi - phi34)) : 0.0;
      ph += nz(cc41) ? cc41 * g5 * std::sin(phi - phi41) : 0.0;
      ph += nz(cc43) ? cc43 * g5 * std::sin(3 * (phi - phi43)) : 0.0;
      ph += nz(cc45) ? cc45 * g5 * std::sin(5 * (phi - phi45)) : 0.0;
      ph += nz(cc52) ? cc52 * g6 * std::sin(2 * (phi - phi52)) : 0.0;
      ph += nz(cc54) ? cc54 * g6 * std::sin(4 * (phi - phi54)) : 0.0;
      ph += nz(cc56) ? cc56 * g6 * std::sin(6 * (phi - phi56)) : 0.0;
    }
    return ph;
  }
};

// ---------------------------------------------------------------------------
// Propagator, gen_propagator_kernel (device_kernels.cuh:215-234) with
// w = -pi lambda dz (propagation_function.cuh:31, :49-57) and the unit
// propagator when dz == 0 (:397-401).
// ---------------------------------------------------------------------------
void propagator(const Grid& g, double lambda, double gx0, double gy0, cd* P) {
  const long long N = g.n();
  if (std::fabs(g.dz) <= 1e-5) {
    for (long long i = 0; i < N; ++i) P[i] = cd(1.0, 0.0);
    return;
  }
  const double w = -kPi * lambda * g.dz;
  for (int iy = 0; iy < g.ny; ++iy)
    for (int ix = 0; ix < g.nx; ++ix) {
      const double gx = g.gx(ix) - gx0, gy = g.gy(iy) - gy0;
      double amp = 1.0 / (double)N;
      if (g.bwl) amp *= g.bwl_factor(ix, iy);
      const double th = w * (gx * gx + gy * gy);
      P[(long long)iy * g.nx + ix] = cd(amp * std::cos(th), amp * std::sin(th));
    }
}

// Generate_probe (device_kernels.cuh:363-388) followed by the unnormalised
// inverse FFT (incident_wave.cuh:29-47).  x_b, y_b: beam position; the phase
// factor is -2 pi (x_b - x_l/2) (grid_dev.cuh:104-111).
void probe(const Grid& g, const Lens& lens, double xb, double yb, cd* psi) {
  const double xe = -2.0 * kPi * (xb - 0.5 * g.lx);
  const double ye = -2.0 * kPi * (yb - 0.5 * g.ly);
  for (int iy = 0; iy < g.ny; ++iy)
    for (int ix = 0; ix < g.nx; ++ix) {
      const double gx = g.gx(ix), gy = g.gy(iy);
      const double g2 = gx * gx + gy * gy;
      cd v(0.0, 0.0);
      if (lens.g2_min <= g2 && g2 < lens.g2_max) {
        const double ph = lens.chi(gx, gy) + xe * gx + ye * gy;
        v = cd(std::cos(ph), std::sin(ph));
      }
      psi[(long long)iy * g.nx + ix] = v;
    }
  fft2(psi, g.nx, g.ny, +1);
}

// ---------------------------------------------------------------------------
// Slicing, Slice_Generator::gene_bounds / get_slicing / fd_by_z
// (slice_generator.hpp:81-168).  The reference instantiates the generator
// only for float (Specimen<float>, specimen.cuh:57, through simulator.cu:7-49),
// so slicing is index work done in float arithmetic: z, dz, eps and R_int are
// rounded to float, the bounds are formed with float operations in the order
// gene_bounds writes them, and every membership test compares floats.
// Direct mode: slice k owns the atoms with z_start <= z < z_end.  Fine mode
// (:88-93, 108-128): bounds cover [z_min - R_int, z_max + R_int] (the
// integration range of Atom_Coord_Data's statistics, in a missing header: the
// build takes R_int = the largest spline radius), and slice k takes the atoms
// with z in [min(z_mid - R_int, z_start), max(z_mid + R_int, z_end)] widened
// by eps on both sides (Inc_Borders).  border: 0 = eMB_Min (STEM/ED/CBED),
// 1 = eMB_Max (HRTEM), 2 = eMB_Middle (specimen.cuh:24-31).  eps =
// epsilon_rel, fixed by the build at 1e-4 (the float tolerance of
// misc_dev.cuh:17-18).  z must be sorted ascending.  Returns the number of
// slices.
// ---------------------------------------------------------------------------
template <typename T>
int slicing_t(int n, const double* zd, double dz_d, int border, double eps_d, bool fine, double r_int_d,
              int max_slices, double* bounds, long long* istart, long long* iend) {
  if (n <= 0) return 0;
  std::vector<T> z(n);
  for (int i = 0; i < n; ++i) z[i] = (T)zd[i];
  const T dz = (T)dz_d, eps = (T)eps_d, r_int = (T)r_int_d;
  const T vmin = fine ? T(z[0] - r_int) : z[0];
  const T vmax = fine ? T(z[n - 1] + r_int) : z[n - 1];
  const T s_v = vmax - vmin;
  const int nv = (int)std::floor(T(s_v + eps) / dz) + 2;
  if (nv - 1 > max_slices) return -(nv - 1);
  std::vector<T> b(nv);
  if (border == 0) {
    const T zb = vmin - T(2) * eps;
    for (int i = 0; i < nv; ++i) b[i] = zb + T(i) * dz;
  } else if (border == 1) {
    const T zt = vmax + T(2) * eps;
    for (int i = 0; i < nv; ++i) b[nv - 1 - i] = zt - T(i) * dz;
  } else {
    const T zb = T(0.5) * (T(vmin + vmax) - T(T(nv - 1) * dz));
    for (int i = 0; i < nv; ++i) b[i] = zb + T(i) * dz;
  }
  for (int i = 0; i < nv; ++i) bounds[i] = b[i];
  for (int s = 0; s < nv - 1; ++s) {
    T zs = b[s], ze = b[s + 1];
    if (fine) {
      const T zm = T(0.5) * (b[s] + b[s + 1]);
      zs = T(std::fmin(T(zm - r_int), b[s]) - eps);
      ze = T(std::fmax(T(zm + r_int), b[s + 1]) + eps);
    }
    long long i0, i1;
    if (zs > ze || z[n - 1] < zs || ze < z[0]) {
      i0 = 1; i1 = 0;
    } else {
      i0 = (zs <= z[0]) ? 0 : (long long)(std::lower_bound(z.begin(), z.end(), zs) - z.begin());
      i1 = (z[n - 1] <= ze) ? (n - 1) : (long long)(std::lower_bound(z.begin(), z.end(), ze) - z.begin()) - 1;
      if (i0 > i1 || z[i1] < zs || ze < z[i0]) { i0 = 1; i1 = 0; }
    }
    istart[s] = i0;
    iend[s] = i1;
  }
  return nv - 1;
}

int slicing(int n, const double* z, double dz, int border, double eps, int max_slices,
            double* bounds, long long* istart, long long* iend) {
  return slicing_t<float>(n, z, dz, border, eps, false, 0.0, max_slices, bounds, istart, iend);
}

int slicing_fine(int n, const double* z, double dz, int border, double eps, double r_int,
                 int max_slices, double* bounds, long long* istart, long long* iend) {
  return slicing_t<float>(n, z, dz, border, eps, true, r_int, max_slices, bounds, istart, iend);
}

// ---------------------------------------------------------------------------
// One slice of projected potential, gen_slice_potential_kernel
// (device_kernels.cuh:28-47) over all atoms of the slice: V += coeff*occ*S(R^2)
// when R^2 < R_max^2, non-periodic distance (grid_dev.cuh:135-154).
// ---------------------------------------------------------------------------
void potential(const Grid& g, long long i0, long long i1, const double* x, const double* y,
               const double* occ, const int* type, int nR, const double* R2, const double* c0,
               const double* c1, const double* c2, const double* c3, const double* rmax2,
               double coeff, double* V) {
  const long long N = g.n();
  for (long long i = 0; i < N; ++i) V[i] = 0.0;
  for (long long a = i0; a <= i1; ++a) {
    const int t = type[a];
    const double* r2 = R2 + (long long)t * nR;
    const double *a0 = c0 + (long long)t * nR, *a1 = c1 + (long long)t * nR,
                 *a2 = c2 + (long long)t * nR, *a3 = c3 + (long long)t * nR;
    const double rm2 = rmax2[t];
    const double rm = std::sqrt(rm2);
    int x0 = std::max(0, (int)std::floor((x[a] - rm) / g.dx) - 1);
    int x1 = std::min(g.nx - 1, (int)std::ceil((x[a] + rm) / g.dx) + 1);
    int y0 = std::max(0, (int)std::floor((y[a] - rm) / g.dy) - 1);
    int y1 = std::min(g.ny - 1, (int)std::ceil((y[a] + rm) / g.dy) + 1);
    for (int iy = y0; iy <= y1; ++iy)
      for (int ix = x0; ix <= x1; ++ix) {
        const double ddx = ix * g.dx - x[a], ddy = iy * g.dy - y[a];
        const double rr = ddx * ddx + ddy * ddy;
        if (rr < rm2)
          V[(long long)iy * g.nx + ix] += coeff * occ[a] * spline_eval(rr, r2, a0, a1, a2, a3, nR);
      }
  }
}

// One fine-sliced slice [zs, ze]: every atom of its range contributes the
// spline of the potential integrated over the slice at its own height
// (projected_potential.cuh:108-116: z_middle = slice centre - atom z).
// Element e of atom a uses pt[e], cl/cnl[e*6 ..], rmax2[e]; splines on nR nodes.
void potential_fine(const Grid& g, double zs, double ze, long long i0, long long i1,
                    const double* x, const double* y, const double* z, const double* occ,
                    const int* type, const int* pt, const double* cl, const double* cnl, int nq,
                    double r_min, double r_max, double tap_frac, int nR, double coeff, double* V) {
  const long long N = g.n();
  for (long long i = 0; i < N; ++i) V[i] = 0.0;
  std::vector<double> R2(nR), c0(nR), c1(nR), c2(nR), c3(nR);
  const double dz = ze - zs, zm = 0.5 * (zs + ze);
  const double rm2 = r_max * r_max, rm = r_max;
  for (long long a = i0; a <= i1; ++a) {
    const int t = type[a];
    fine_spline(pt[t], cl + 6 * t, cnl + 6 * t, zm - z[a], dz, nq, r_min, r_max, tap_frac, nR,
                R2.data(), c0.data(), c1.data(), c2.data(), c3.data());
    int x0 = std::max(0, (int)std::floor((x[a] - rm) / g.dx) - 1);
    int x1 = std::min(g.nx - 1, (int)std::ceil((x[a] + rm) / g.dx) + 1);
    int y0 = std::max(0, (int)std::floor((y[a] - rm) / g.dy) - 1);
    int y1 = std::min(g.ny - 1, (int)std::ceil((y[a] + rm) / g.dy) + 1);
    for (int iy = y0; iy <= y1; ++iy)
      for (int ix = x0; ix <= x1; ++ix) {
        const double ddx = ix * g.dx - x[a], ddy = iy * g.dy - y[a];
        const double rr = ddx * ddx + ddy * ddy;
        if (rr < rm2)
          V[(long long)iy * g.nx + ix] +=
              coeff * occ[a] * spline_eval(rr, R2.data(), c0.data(), c1.data(), c2.data(), c3.data(), nR);
      }
  }
}

// ---------------------------------------------------------------------------
// Multislice sequence, Wave_Function::wave_modulation (wave_function.cuh:60-73):
// for k < S-1: psi <- IFFT[P . FFT[t_k . psi]]; then Psi = FFT[t_{S-1} . psi],
// left in Fourier space and unnormalised.  t_k = exp(i V_k) with V_k already
// scaled by sigma (transmission_function.cuh:107-111, fftcallbacks.cuh:36-49),
// or a given complex transmission when tc != nullptr.  If ckpt != nullptr,
// the wave entering slice k (real space) is stored in ckpt[k].
// ---------------------------------------------------------------------------
template <class R>
void multislice_t(const Grid& g, const std::complex<R>* P, int S, const R* V, const std::complex<R>* tc,
                  std::complex<R>* psi, std::complex<R>* ckpt) {
  typedef std::complex<R> C;
  const long long N = g.n();
  for (int k = 0; k < S; ++k) {
    if (ckpt) std::copy(psi, psi + N, ckpt + (long long)k * N);
    for (long long i = 0; i < N; ++i) {
      C t;
      if (tc) t = tc[(long long)k * N + i];
      else {
        const R v = V[(long long)k * N + i];
        t = C(std::cos(v), std::sin(v));
      }
      psi[i] *= t;
    }
    fft2(psi, g.nx, g.ny, -1);
    if (k == S - 1) break;
    for (long long i = 0; i < N; ++i) psi[i] *= P[i];
    fft2(psi, g.nx, g.ny, +1);
  }
}
void multislice(const Grid& g, const cd* P, int S, const double* V, const cd* tc, cd* psi, cd* ckpt) {
  multislice_t<double>(g, P, S, V, tc, psi, ckpt);
}

// sum_square_over_Det (device_kernels.cuh:534-579): closed band g2_min <= g^2 <= g2_max.
template <class R>
double detector_sum(const Grid& g, const std::complex<R>* Psi, double g2min, double g2max) {
  double s = 0.0;
  for (int iy = 0; iy < g.ny; ++iy)
    for (int ix = 0; ix < g.nx; ++ix) {
      const double g2 = g.gx(ix) * g.gx(ix) + g.gy(iy) * g.gy(iy);
      if (g2min <= g2 && g2 <= g2max) s += (double)std::norm(Psi[(long long)iy * g.nx + ix]);
    }
  return s;
}

}  // namespace orc

using namespace orc;

// ===========================================================================
// C ABI for ctypes (tests and the CPU baseline).  All arrays are host arrays.
// ===========================================================================
// ---------------------------------------------------------------------------
// Partial coherence of the objective lens (Lens_Modulation::PCTF_TEM /
// TCC_TEM, lens_modulation.cuh:91-222; apply_PCTF_kernel,
// device_kernels.cuh:504-518; incoherent_phase_shift,
// basic_functions_dev.cuh:236-266).
//
// Conventions fixed by the build (the reference's constants live in missing
// headers): sigma = standard deviation of the Gaussian defocus spread (A);
// beta = e^-1 half-width semi-angle of the Gaussian source (rad), i.e. tilt
// density ~ exp(-|q|^2 / q0^2) with q0 = sin(beta) / lambda.
//   PCTF (quasi-coherent, linear imaging): Psi(g) A(g) E(g) e^{i chi(g)} with
//     u = 1 + 2 (pi beta sigma)^2 g^2,
//     E = exp(-[(pi beta)^2 g^2 (c30 lambda^2 g^2 + c10)^2
//              + (pi lambda sigma g^2)^2 / 2] / u) / sqrt(u),
//   which is the exact Gaussian average of the two-beam interference term to
//   first order in the tilt (the reference writes (c30 lambda^2 g^2 - c10),
//   the opposite defocus sign to its own chi, and its defocus width as the
//   e^-1 half-width; both are made consistent with the TCC here).
//   TCC: sum over Gauss-Hermite nodes of coherent CTF images,
//     defocus c10 + sqrt(2) sigma t_j, weight w_j / sqrt(pi);
//     tilt q = q0 (t_i, t_k), weight w_i w_k / pi;
//   mode 1 = both, 2 = temporal only, 3 = spatial only.
// ---------------------------------------------------------------------------
// Gauss-Hermite nodes / weights for weight exp(-t^2): Newton on the
// orthonormal Hermite recurrence (textbook), ascending nodes.
void gauss_hermite(int n, double* x, double* w) {
  const double pim4 = 0.7511255444649425;  // pi^(-1/4)
  const int m = (n + 1) / 2;
  double z = 0.0;
  for (int i = 0; i < m; ++i) {
    if (i == 0) z = std::sqrt(2.0 * n + 1.0) - 1.85575 * std::pow(2.0 * n + 1.0, -0.16667);
    else if (i == 1) z -= 1.14 * std::pow((double)n, 0.426) / z;
    else if (i == 2) z = 1.86 * z - 0.86 * x[0];
    else if (i == 3) z = 1.91 * z - 0.91 * x[1];
    else z = 2.0 * z - x[i - 2];
    double pp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = pim4, p2 = 0.0;
      for (int j = 0; j < n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = z * std::sqrt(2.0 / (j + 1)) * p2 - std::sqrt((double)j / (j + 1)) * p3;
      }
      pp = std::sqrt(2.0 * n) * p2;
      const double dz = p1 / pp;
      z -= dz;
      if (std::fabs(dz) < 1e-15) break;
    }
    x[i] = z;  // descending positive nodes for now
    w[i] = 2.0 / (pp * pp);
  }
  // mirror into ascending order
  std::vector<double> xs(n), ws(n);
  for (int i = 0; i < m; ++i) {
    xs[n - 1 - i] = x[i];
    ws[n - 1 - i] = w[i];
    xs[i] = -x[i];
    ws[i] = w[i];
  }
  if (n % 2 == 1) xs[n / 2] = 0.0;
  for (int i = 0; i < n; ++i) {
    x[i] = xs[i];
    w[i] = ws[i];
  }
}

double pctf_envelope(double g2, double lambda, double c10, double c30, double sigma, double beta) {
  const double c = kPi * beta * sigma;
  const double u = 1.0 + 2.0 * c * c * g2;
  const double tc = kPi * lambda * sigma * g2;
  const double temp = 0.5 * tc * tc;
  const double sc = kPi * beta * (c30 * lambda * lambda * g2 + c10);
  const double spa = sc * sc * g2;
  return std::exp(-(spa + temp) / u) / std::sqrt(u);
}

// Coherent CTF image with the lens shifted by the tilt (gxu, gyu) and the
// defocus by dc10 (apply_CTF_kernel, device_kernels.cuh:445-468), |.|^2 or
// (re, im) scaled by wt and accumulated.
static void ctf_accumulate(const Grid& g, const double* lens, double lambda, double dc10, double gxu,
                           double gyu, const cd* in, double wt, double* inten, double* re, double* im) {
  std::vector<double> lp(lens, lens + L_COUNT);
  lp[L_C10] += dc10;
  Lens l(lp.data(), lambda, 0.0);
  const long long N = g.n();
  std::vector<cd> w(N);
  for (int iy = 0; iy < g.ny; ++iy)
    for (int ix = 0; ix < g.nx; ++ix) {
      const double gx = g.gx(ix) + gxu, gy = g.gy(iy) + gyu, g2 = gx * gx + gy * gy;
      const long long i = (long long)iy * g.nx + ix;
      if (l.g2_min <= g2 && g2 < l.g2_max) {
        const double ph = l.chi(gx, gy);
        w[i] = in[i] * cd(std::cos(ph), std::sin(ph));
      } else {
        w[i] = 0.0;
      }
    }
  fft2(w.data(), g.nx, g.ny, +1);
  for (long long i = 0; i < N; ++i) {
    if (inten) inten[i] += wt * std::norm(w[i]);
    if (re) re[i] += wt * w[i].real();
    if (im) im[i] += wt * w[i].imag();
  }
}

extern "C" {

int orc_version(void) { return 1; }

void orc_gauss_hermite(int n, double* x, double* w) { gauss_hermite(n, x, w); }

// PCTF image (PCTF_TEM): sigma (A), beta (mrad), mode as above.
void orc_pctf_image(int nx, int ny, double lx, double ly, double lambda, const double* lens, double sigma,
                    double beta_mrad, int mode, const double* Psi, double* inten, double* re, double* im) {
  Grid g(nx, ny, lx, ly, 0.0, 0);
  Lens l(lens, lambda, 0.0);
  const double s = (mode == 3) ? 0.0 : sigma, b = (mode == 2) ? 0.0 : beta_mrad * 1e-3;
  const long long N = g.n();
  std::vector<cd> w(N);
  const cd* in = reinterpret_cast<const cd*>(Psi);
  for (int iy = 0; iy < ny; ++iy)
    for (int ix = 0; ix < nx; ++ix) {
      const double gx = g.gx(ix), gy = g.gy(iy), g2 = gx * gx + gy * gy;
      const long long i = (long long)iy * nx + ix;
      if (l.g2_min <= g2 && g2 < l.g2_max) {
        const double ph = l.chi(gx, gy);
        const double e = pctf_envelope(g2, lambda, lens[L_C10], lens[L_C30], s, b);
        w[i] = in[i] * std::polar(e, ph);
      } else {
        w[i] = 0.0;
      }
    }
  fft2(w.data(), nx, ny, +1);
  for (long long i = 0; i < N; ++i) {
    if (inten) inten[i] = std::norm(w[i]);
    if (re) re[i] = w[i].real();
    if (im) im[i] = w[i].imag();
  }
}

// TCC image (TCC_TEM): nt / ns Gauss-Hermite points of the defocus and of
// each tilt axis.
void orc_tcc_image(int nx, int ny, double lx, double ly, double lambda, const double* lens, double sigma,
                   int nt, double beta_mrad, int ns, int mode, const double* Psi, double* inten, double* re,
                   double* im) {
  Grid g(nx, ny, lx, ly, 0.0, 0);
  const long long N = g.n();
  for (long long i = 0; i < N; ++i) {
    if (inten) inten[i] = 0.0;
    if (re) re[i] = 0.0;
    if (im) im[i] = 0.0;
  }
  const cd* in = reinterpret_cast<const cd*>(Psi);
  std::vector<double> tx(std::max(nt, 1)), tw(std::max(nt, 1)), sx(std::max(ns, 1)), sw(std::max(ns, 1));
  const bool temporal = mode != 3, spatial = mode != 2;
  const int NT = temporal ? nt : 1, NS = spatial ? ns : 1;
  if (temporal) gauss_hermite(nt, tx.data(), tw.data()); else { tx[0] = 0.0; tw[0] = std::sqrt(kPi); }
  if (spatial) gauss_hermite(ns, sx.data(), sw.data()); else { sx[0] = 0.0; sw[0] = std::sqrt(kPi); }
  const double q0 = std::sin(beta_mrad * 1e-3) / lambda;
  for (int a = 0; a < NS; ++a)
    for (int b = 0; b < NS; ++b) {
      if (!spatial && (a > 0 || b > 0)) continue;
      const double ws = spatial ? sw[a] * sw[b] / kPi : 1.0;
      const double qx = spatial ? q0 * sx[a] : 0.0, qy = spatial ? q0 * sx[b] : 0.0;
      for (int j = 0; j < NT; ++j) {
        const double wt = temporal ? tw[j] / std::sqrt(kPi) : 1.0;
        const double df = temporal ? std::sqrt(2.0) * sigma * tx[j] : 0.0;
        ctf_accumulate(g, lens, lambda, df, qx, qy, in, ws * wt, inten, re, im);
      }
    }
}

void orc_constants(double e_kv, double* lam, double* gam, double* sig) {
  *lam = wavelength(e_kv);
  *gam = gamma_factor(e_kv);
  *sig = sigma_factor(e_kv);
}

void orc_bessel_k01(double x, double* k0, double* k1) { bessel_k01(x, k0, k1); }

void orc_moliere_coefs(int Z, double* a, double* b) { moliere_coefs(Z, a, b); }

void orc_vr(int nk, const double* a, const double* b, int ng, const double* c,
            const double* d, int n, const double* R, double* V, double* dV) {
  for (int i = 0; i < n; ++i) vr_kirkland(nk, a, b, ng, c, d, R[i], V + i, dV + i);
}

void orc_build_spline(int nk, const double* a, const double* b, int ng, const double* c,
                      const double* d, double r_min, double r_max, double tap_frac, int nR,
                      double* R2, double* c0, double* c1, double* c2, double* c3) {
  build_spline(nk, a, b, ng, c, d, r_min, r_max, tap_frac, nR, R2, c0, c1, c2, c3);
}

void orc_spline_eval(int n, const double* x, int nR, const double* R2, const double* c0,
                     const double* c1, const double* c2, const double* c3, double* out) {
  for (int i = 0; i < n; ++i) out[i] = spline_eval(x[i], R2, c0, c1, c2, c3, nR);
}

void orc_fft1d(double* data, int n, int sign) {
  Fft f(n);
  std::vector<cd> bi(n), bo(n), tmp(n);
  f.run(reinterpret_cast<cd*>(data), 1, sign, bi.data(), bo.data(), tmp.data());
}

void orc_fft2d(double* data, int nx, int ny, int sign) {
  fft2(reinterpret_cast<cd*>(data), nx, ny, sign);
}

int orc_slicing(int n, const double* z, double dz, int border, double eps, int max_slices,
                double* bounds, long long* istart, long long* iend) {
  return slicing(n, z, dz, border, eps, max_slices, bounds, istart, iend);
}

int orc_slicing_fine(int n, const double* z, double dz, int border, double eps, double r_int,
                     int max_slices, double* bounds, long long* istart, long long* iend) {
  return slicing_fine(n, z, dz, border, eps, r_int, max_slices, bounds, istart, iend);
}

void orc_vr3d(int pt, int n, const double* r, const double* cl, const double* cnl, double* V,
              double* dVrir) {
  for (int i = 0; i < n; ++i) vr3d(pt, r[i], cl, cnl, 1.0, V + i, dVrir + i);
}

void orc_gauss_legendre(int n, double* x, double* w) { gauss_legendre(n, x, w); }

void orc_moliere_3d(int Z, double* cl, double* cnl) { moliere_3d(Z, cl, cnl); }

void orc_fine_spline(int pt, const double* cl, const double* cnl, double a, double dz, int nq,
                     double r_min, double r_max, double tap_frac, int nR, double* R2, double* c0,
                     double* c1, double* c2, double* c3) {
  fine_spline(pt, cl, cnl, a, dz, nq, r_min, r_max, tap_frac, nR, R2, c0, c1, c2, c3);
}

void orc_potential_fine(int nx, int ny, double lx, double ly, double zs, double ze, long long i0,
                        long long i1, const double* x, const double* y, const double* z,
                        const double* occ, const int* type, const int* pt, const double* cl,
                        const double* cnl, int nq, double r_min, double r_max, double tap_frac,
                        int nR, double coeff, double* V) {
  Grid g(nx, ny, lx, ly, 0.0, 0);
  potential_fine(g, zs, ze, i0, i1, x, y, z, occ, type, pt, cl, cnl, nq, r_min, r_max, tap_frac, nR,
                 coeff, V);
}

void orc_potential(int nx, int ny, double lx, double ly, long long i0, long long i1,
                   const double* x, const double* y, const double* occ, const int* type, int nR,
                   const double* R2, const double* c0, const double* c1, const double* c2,
                   const double* c3, const double* rmax2, double coeff, double* V) {
  Grid g(nx, ny, lx, ly, 0.0, 0);
  potential(g, i0, i1, x, y, occ, type, nR, R2, c0, c1, c2, c3, rmax2, coeff, V);
}

void orc_propagator(int nx, int ny, double lx, double ly, double dz, int bwl, double lambda,
                    double gx0, double gy0, double* P) {
  Grid g(nx, ny, lx, ly, dz, bwl);
  propagator(g, lambda, gx0, gy0, reinterpret_cast<cd*>(P));
}

void orc_bwl_params(int nx, int ny, double lx, double ly, double* gl2_max, double* alpha) {
  Grid g(nx, ny, lx, ly, 0.0, 1);
  *gl2_max = g.gl2_max;
  *alpha = g.alpha;
}

// lens: L_COUNT doubles in the em_lens field order
void orc_lens_g2(const double* lens, double lambda, double* g2min, double* g2max) {
  Lens l(lens, lambda, 0.0);
  *g2min = l.g2_min;
  *g2max = l.g2_max;
}

void orc_chi(const double* lens, double lambda, double defocus_shift, int n, const double* gx,
             const double* gy, double* out) {
  Lens l(lens, lambda, defocus_shift);
  for (int i = 0; i < n; ++i) out[i] = l.chi(gx[i], gy[i]);
}

// Probe in real space: Generate_probe + inverse FFT; defocus_shift is z_mid of
// slice 0 (incident_wave.cuh:33-35).
void orc_probe(int nx, int ny, double lx, double ly, double lambda, const double* lens,
               double defocus_shift, double xb, double yb, double* psi) {
  Grid g(nx, ny, lx, ly, 0.0, 0);
  Lens l(lens, lambda, defocus_shift);
  probe(g, l, xb, yb, reinterpret_cast<cd*>(psi));
}

// psi: in = incident wave (real space), out = exit wave (Fourier, unnormalised).
// V: [S][N] sigma-scaled potential (or null), tc: [S][N] complex transmission (or null).
void orc_multislice(int nx, int ny, const double* P, int S, const double* V, const double* tc,
                    double* psi, double* ckpt) {
  Grid g(nx, ny, 1.0, 1.0, 0.0, 0);
  multislice(g, reinterpret_cast<const cd*>(P), S, V, reinterpret_cast<const cd*>(tc),
             reinterpret_cast<cd*>(psi), reinterpret_cast<cd*>(ckpt));
}

double orc_detector_sum(int nx, int ny, double lx, double ly, const double* Psi, double g2min,
                        double g2max) {
  Grid g(nx, ny, lx, ly, 0.0, 0);
  return detector_sum(g, reinterpret_cast<const cd*>(Psi), g2min, g2max);
}

// Coherent CTF image, Lens_Modulation::CTF_TEM (lens_modulation.cuh:78-89) with
// apply_CTF_kernel (device_kernels.cuh:445-468): IFFT_u[Psi e^{i chi} A] then
// |.|^2 (intensity) and/or re/im (any output pointer may be null).
void orc_ctf_image(int nx, int ny, double lx, double ly, double lambda, const double* lens,
                   const double* Psi, double* inten, double* re, double* im) {
  Grid g(nx, ny, lx, ly, 0.0, 0);
  Lens l(lens, lambda, 0.0);
  const long long N = g.n();
  std::vector<cd> w(N);
  const cd* in = reinterpret_cast<const cd*>(Psi);
  for (int iy = 0; iy < ny; ++iy)
    for (int ix = 0; ix < nx; ++ix) {
      const double gx = g.gx(ix), gy = g.gy(iy), g2 = gx * gx + gy * gy;
      const long long i = (long long)iy * nx + ix;
      if (l.g2_min <= g2 && g2 < l.g2_max) {
        const double ph = l.chi(gx, gy);
        w[i] = in[i] * cd(std::cos(ph), std::sin(ph));
      } else {
        w[i] = 0.0;
      }
    }
  fft2(w.data(), nx, ny, +1);
  for (long long i = 0; i < N; ++i) {
    if (inten) inten[i] = std::norm(w[i]);
    if (re) re[i] = w[i].real();
    if (im) im[i] = w[i].imag();
  }
}

// Thickness series, run_thickness_simu (thickness_simulator.cuh:11-85) with
// the full-wave copy the reference intends (its :38-39 copies half): after
// every slice (single_slice_modulation = transmit + propagate, wave_function.cuh:88-94)
// band-pass the wave by [g2_min, g2_max] with 1/N (Filter_wave,
// device_kernels.cuh:389-410, wave_function.cuh:163-174) and record |.|^2.
// frames: [S][N]; only slices with (k+1) % stride == 0 or k == S-1 are
// recorded when stride > 1 (frame index = count of recorded slices so far).
int orc_thickness(int nx, int ny, double lx, double ly, const double* P, int S, const double* V,
                  double* psi, double g2min, double g2max, int stride, double* frames) {
  Grid g(nx, ny, lx, ly, 0.0, 0);
  const long long N = g.n();
  cd* w = reinterpret_cast<cd*>(psi);
  const cd* Pc = reinterpret_cast<const cd*>(P);
  std::vector<cd> tmp(N);
  int nf = 0;
  for (int k = 0; k < S; ++k) {
    for (long long i = 0; i < N; ++i) {
      const double v = V[(long long)k * N + i];
      w[i] *= cd(std::cos(v), std::sin(v));
    }
    fft2(w, nx, ny, -1);
    for (long long i = 0; i < N; ++i) w[i] *= Pc[i];
    fft2(w, nx, ny, +1);
    const bool rec = stride <= 1 || ((k + 1) % stride == 0) || k == S - 1;
    if (!rec) continue;
    std::copy(w, w + N, tmp.begin());
    fft2(tmp.data(), nx, ny, -1);
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const double gg = g.gx(ix) * g.gx(ix) + g.gy(iy) * g.gy(iy);
        const long long i = (long long)iy * nx + ix;
        tmp[i] = (g2min <= gg && gg <= g2max) ? tmp[i] / (double)N : cd(0.0, 0.0);
      }
    fft2(tmp.data(), nx, ny, +1);
    for (long long i = 0; i < N; ++i) frames[(long long)nf * N + i] = std::norm(tmp[i]);
    ++nf;
  }
  return nf;
}

// Batched STEM oracle, run_STEM_tile_simu's probe loop (scan_simulator.cuh:87-109)
// with the batched probe builder (Generate_multi_probe, device_kernels.cuh:411-444):
// for each probe b: build the probe at (xb[b], yb[b]), run the multislice over
// the stored sigma-scaled potentials V[S][N], then the weighted detector sums
// out[d*B + b] = w * sum |Psi|^2 over [g2min[d], g2max[d]].  OpenMP over
// probes, one serial FFT per thread.  Returns the number of threads used.
int orc_stem(int nx, int ny, double lx, double ly, double lambda, const double* lens,
             double defocus_shift, const double* P, int S, const double* V, int B,
             const double* xb, const double* yb, int ndet, const double* g2min,
             const double* g2max, double w, double* out, int nthreads) {
  Grid g(nx, ny, lx, ly, 0.0, 0);
  Lens l(lens, lambda, defocus_shift);
  const long long N = g.n();
  const cd* Pc = reinterpret_cast<const cd*>(P);
  int used = 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
  {
#pragma omp single
    used = omp_get_num_threads();
    std::vector<cd> psi(N);
#pragma omp for schedule(dynamic, 1)
    for (int b = 0; b < B; ++b) {
      probe(g, l, xb[b], yb[b], psi.data());
      multislice(g, Pc, S, V, nullptr, psi.data(), nullptr);
      for (int d = 0; d < ndet; ++d) out[(long long)d * B + b] = w * detector_sum(g, psi.data(), g2min[d], g2max[d]);
    }
  }
#else
  (void)nthreads;
  std::vector<cd> psi(N);
  for (int b = 0; b < B; ++b) {
    probe(g, l, xb[b], yb[b], psi.data());
    multislice(g, Pc, S, V, nullptr, psi.data(), nullptr);
    for (int d = 0; d < ndet; ++d) out[(long long)d * B + b] = w * detector_sum(g, psi.data(), g2min[d], g2max[d]);
  }
#endif
  return used;
}

// The same STEM loop with the wave, transmission, propagator and transforms
// in FP32 (probe built in FP64 and rounded; detector sums accumulated in
// FP64): the FP32 variant of the CPU baseline, timed beside the FP64 oracle.
int orc_stem_f32(int nx, int ny, double lx, double ly, double lambda, const double* lens,
                 double defocus_shift, const double* P, int S, const double* V, int B,
                 const double* xb, const double* yb, int ndet, const double* g2min,
                 const double* g2max, double w, double* out, int nthreads) {
  typedef std::complex<float> cf;
  Grid g(nx, ny, lx, ly, 0.0, 0);
  Lens l(lens, lambda, defocus_shift);
  const long long N = g.n();
  std::vector<cf> Pf(N);
  std::vector<float> Vf((size_t)S * N);
  for (long long i = 0; i < N; ++i) Pf[i] = cf((float)P[2 * i], (float)P[2 * i + 1]);
  for (long long i = 0; i < (long long)S * N; ++i) Vf[i] = (float)V[i];
  int used = 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
  {
#pragma omp single
    used = omp_get_num_threads();
    std::vector<cd> p64(N);
    std::vector<cf> psi(N);
#pragma omp for schedule(dynamic, 1)
    for (int b = 0; b < B; ++b) {
      probe(g, l, xb[b], yb[b], p64.data());
      for (long long i = 0; i < N; ++i) psi[i] = cf((float)p64[i].real(), (float)p64[i].imag());
      multislice_t<float>(g, Pf.data(), S, Vf.data(), nullptr, psi.data(), nullptr);
      for (int d = 0; d < ndet; ++d) out[(long long)d * B + b] = w * detector_sum(g, psi.data(), g2min[d], g2max[d]);
    }
  }
#else
  (void)nthreads;
  std::vector<cd> p64(N);
  std::vector<cf> psi(N);
  for (int b = 0; b < B; ++b) {
    probe(g, l, xb[b], yb[b], p64.data());
    for (long long i = 0; i < N; ++i) psi[i] = cf((float)p64[i].real(), (float)p64[i].imag());
    multislice_t<float>(g, Pf.data(), S, Vf.data(), nullptr, psi.data(), nullptr);
    for (int d = 0; d < ndet; ++d) out[(long long)d * B + b] = w * detector_sum(g, psi.data(), g2min[d], g2max[d]);
  }
#endif
  return used;
}

// Adjoint of the multislice with respect to the per-slice complex
// transmission t[S][N] (no reference counterpart: the reference has no
// adjoint).  Forward as orc_multislice with tc = t for each of the B probes
// psi0[B][N]; given gPsi[B][N] = dL/dRe(Psi) + i dL/dIm(Psi) for the exit
// waves, returns gt[S][N] = dL/dRe(t) + i dL/dIm(t) summed over probes:
//   lambda_{S-1} = FFT^H gPsi;  gt_k += conj(psi_k) lambda_k;
//   lambda_{k-1} = FFT^H conj(P) FFT [conj(t_k) lambda_k],
// with FFT^H the unnormalised inverse transform.
void orc_adjoint(int nx, int ny, const double* P, int S, const double* t, int B,
                 const double* psi0, const double* gPsi, double* gt, double* Psi_out) {
  Grid g(nx, ny, 1.0, 1.0, 0.0, 0);
  const long long N = g.n();
  const cd* Pc = reinterpret_cast<const cd*>(P);
  const cd* tc = reinterpret_cast<const cd*>(t);
  cd* gtc = reinterpret_cast<cd*>(gt);
  for (long long i = 0; i < (long long)S * N; ++i) gtc[i] = 0.0;
  std::vector<cd> ck((long long)S * N), psi(N), lam(N);
  for (int b = 0; b < B; ++b) {
    const cd* p0 = reinterpret_cast<const cd*>(psi0) + (long long)b * N;
    std::copy(p0, p0 + N, psi.begin());
    multislice(g, Pc, S, nullptr, tc, psi.data(), ck.data());
    if (Psi_out) std::copy(psi.begin(), psi.end(), reinterpret_cast<cd*>(Psi_out) + (long long)b * N);
    const cd* gp = reinterpret_cast<const cd*>(gPsi) + (long long)b * N;
    std::copy(gp, gp + N, lam.begin());
    fft2(lam.data(), nx, ny, +1);
    for (int k = S - 1; k >= 0; --k) {
      const cd* pk = ck.data() + (long long)k * N;
      const cd* tk = tc + (long long)k * N;
      cd* gk = gtc + (long long)k * N;
      for (long long i = 0; i < N; ++i) {
        gk[i] += std::conj(pk[i]) * lam[i];
        lam[i] *= std::conj(tk[i]);
      }
      if (k == 0) break;
      fft2(lam.data(), nx, ny, -1);
      for (long long i = 0; i < N; ++i) lam[i] *= std::conj(Pc[i]);
      fft2(lam.data(), nx, ny, +1);
    }
  }
}

}  // extern "C"
