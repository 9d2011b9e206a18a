// The engine's host routines driven over their parameter ranges, built with
// AddressSanitizer and UBSan by tests/test_host_sanitized.py: slicing (direct
// and fine), splines for every element, quadratures, lens and grid derivation.
#include <cstdio>
#include <random>
#include <algorithm>
#include "em_host.hpp"
using namespace em;
int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  long long checks = 0;
  for (int t = 0; t < 400; ++t) {
    const int n = (int)(U(rng) * 300);
    std::vector<double> z(n);
    const double span = 0.5 + U(rng) * 80;
    for (auto& v : z) v = U(rng) * span - 3.0;
    std::sort(z.begin(), z.end());
    const double dz = 0.3 + U(rng) * 3;
    for (int border = 0; border < 3; ++border) {
      auto s = make_slices(z, dz, border, 1e-4);
      auto f = make_slices_fine(z, dz, border, 1e-4, U(rng) * 4);
      for (auto& r : s) if (r.i0 <= r.i1 && (r.i0 < 0 || r.i1 >= n)) { std::printf("bad range\n"); return 1; }
      for (auto& r : f) if (r.i0 <= r.i1 && (r.i0 < 0 || r.i1 >= n)) { std::printf("bad fine range\n"); return 1; }
      checks += s.size() + f.size();
    }
  }
  for (int Z = 1; Z <= 103; ++Z)
    for (int nR : {4, 16, 128, 256}) {
      Spline sp = build_spline(moliere_coefs(Z), nR, 0.005, 3.0, 0.85);
      checks += sp.R2.size();
      double cl[6], cnl[6];
      moliere_3d(Z, cl, cnl);
    }
  std::vector<double> x, w;
  for (int n = 1; n <= 1024; n += (n < 64 ? 1 : 37)) { gauss_legendre(n, x, w); checks += x.size(); }
  for (int n = 1; n <= 64; ++n) { gauss_hermite(n, x, w); checks += x.size(); }
  em_lens l{};
  l.c_10 = 50; l.c_30 = 1e7; l.outer_aper_mrad = 25; l.c_12 = 3; l.phi_12 = 0.4;
  for (double kv : {80.0, 200.0, 300.0}) {
    Constants k = beam_constants(kv);
    LensDev d = derive_lens(l, k.lambda, 12.5);
    checks += (d.g2_max > 0) + (mrad_to_g2(25, k.lambda, true) > 0) + (mrad_to_g2(0, k.lambda, true) > 0);
  }
  em_grid g{256, 128, 20.0, 10.0, 1.0, 1};
  GridDev gd = make_grid_dev(g);
  checks += gd.nx;
  std::printf("host checks %lld done\n", checks);
  return 0;
}
