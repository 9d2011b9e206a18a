// em_simulators.cpp -- the reference's C++ simulator surface
// (include/cudaEM/simulator.hpp) on top of the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "cudaEM/simulator.hpp"
#include "cudaem_amd.h"
#include "em_kernels.hpp"

namespace cudaEM {
namespace {

void ok(em_status s, em_plan p, const char* what) {
  if (s != EM_OK)
    throw simulation_error(std::string(what) + " failed (status " + std::to_string(s) + "): " +
                           (p ? em_plan_last_error(p) : ""));
}
void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw simulation_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct Dev {
  void* p = nullptr;
  explicit Dev(size_t bytes) { hip_ok(hipMalloc(&p, std::max<size_t>(bytes, 16)), "hipMalloc"); }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  template <class U>
  U* as() const {
    return static_cast<U*>(p);
  }
};

struct Plan {
  em_plan p = nullptr;
  Plan(const em_grid& g, const em_beam& b, int batch) { ok(em_plan_create(&g, &b, batch, &p), nullptr, "em_plan_create"); }
  ~Plan() {
    if (p) em_plan_destroy(p);
  }
};

// Potential options of a run: fine slicing and the elements' 3D coefficients.
template <class T>
void configure_potential(em_plan p, const multislice_params<T>& m) {
  for (const Element_Potential<T>& e : m.element_potentials) {
    double cl[6], cnl[6];
    for (int i = 0; i < 6; ++i) {
      cl[i] = e.cl[i];
      cnl[i] = e.cnl[i];
    }
    ok(em_set_element_params(p, e.Z, (int)m.potential_type, cl, cnl), p, "em_set_element_params");
  }
  ok(em_plan_set_fine_slicing(p, m.enable_fine_slicing ? 1 : 0, m.fine_slicing_nq), p, "em_plan_set_fine_slicing");
}

template <class T>
em_grid grid_of(const multislice_params<T>& m) {
  em_grid g;
  g.nx = m.nx;
  g.ny = m.ny;
  g.lx = m.lx;
  g.ly = m.ly;
  g.dz = m.dz;
  g.bwl = m.bwl ? 1 : 0;
  return g;
}

template <class T>
em_beam beam_of(const multislice_params<T>& m) {
  em_beam b;
  b.e0_kv = m.E_0;
  b.theta = m.theta;
  b.phi = m.phi;
  return b;
}

template <class T>
em_lens lens_of(const Lens<T>& l) {
  em_lens o;
  o.c_10 = l.c_10; o.c_12 = l.c_12; o.phi_12 = l.phi_12; o.c_21 = l.c_21; o.phi_21 = l.phi_21;
  o.c_23 = l.c_23; o.phi_23 = l.phi_23; o.c_30 = l.c_30; o.c_32 = l.c_32; o.phi_32 = l.phi_32;
  o.c_34 = l.c_34; o.phi_34 = l.phi_34; o.c_41 = l.c_41; o.phi_41 = l.phi_41; o.c_43 = l.c_43;
  o.phi_43 = l.phi_43; o.c_45 = l.c_45; o.phi_45 = l.phi_45; o.c_50 = l.c_50; o.c_52 = l.c_52;
  o.phi_52 = l.phi_52; o.c_54 = l.c_54; o.phi_54 = l.phi_54; o.c_56 = l.c_56; o.phi_56 = l.phi_56;
  o.m = l.m;
  o.inner_aper_mrad = l.inner_aper_ang;
  o.outer_aper_mrad = l.outer_aper_ang;
  return o;
}

template <class T>
int set_specimen(em_plan p, const Atom_Coord_Data<T>& a, int border) {
  const size_t n = a.size();
  if (a.y.size() != n || a.z.size() != n || a.Z.size() != n || (!a.occ.empty() && a.occ.size() != n))
    throw simulation_error("atom arrays of unequal length");
  std::vector<double> x(a.x.begin(), a.x.end()), y(a.y.begin(), a.y.end()), z(a.z.begin(), a.z.end());
  std::vector<double> occ(n, 1.0);
  if (!a.occ.empty()) occ.assign(a.occ.begin(), a.occ.end());
  em_atoms at;
  at.n = (long long)n;
  at.x = x.data();
  at.y = y.data();
  at.z = z.data();
  at.Z = a.Z.data();
  at.occ = occ.data();
  int S = 0;
  ok(em_plan_set_specimen(p, &at, border, &S), p, "em_plan_set_specimen");
  return S;
}

double zmid0(em_plan p) {
  double zs = 0, ze = 0;
  ok(em_slice_info(p, 0, &zs, &ze, nullptr, nullptr), p, "em_slice_info");
  return 0.5 * (zs + ze);
}

template <class T>
std::vector<double> det_angles(const std::vector<T>& v) {
  return std::vector<double>(v.begin(), v.end());
}

// Frozen-phonon configurations of a run and their weight (get_weight,
// multislice_params_dev.cuh:209-216; all configurations on this device).
template <class T>
void phonon_range(const multislice_params<T>& m, int* n, float* w) {
  if (m.frozen_phonon_enable()) {
    *n = std::max(1, m.pn_nconf);
    *w = 1.f / (float)*n;
  } else {
    *n = 1;
    *w = 1.f;
  }
}

template <class T>
Atom_Coord_Data<T> atoms_of_config(const multislice_params<T>& m, int iconf) {
  if (!m.frozen_phonon_enable()) return m.atoms;
  Atom_Coord_Data<T> out;
  vibrate_atom_positions(m.atoms, out, m.pn_seed, iconf, m.pn_sigma, m.pn_dim);
  return out;
}

template <class T>
void scan_positions(const multislice_params<T>& m, std::vector<double>& x, std::vector<double>& y) {
  const double x0 = m.scan_range[0] * m.lx, y0 = m.scan_range[1] * m.ly;
  const double wx = (m.scan_range[2] - m.scan_range[0]) * m.lx, wy = (m.scan_range[3] - m.scan_range[1]) * m.ly;
  x.clear();
  y.clear();
  for (int iy = 0; iy < m.STEM_row; ++iy)
    for (int ix = 0; ix < m.STEM_col; ++ix) {
      x.push_back(x0 + ix * wx / m.STEM_col);
      y.push_back(y0 + iy * wy / m.STEM_row);
    }
}

// Incident intensity sum_g |Psi_0(g)|^2 and the unscattered probe on each
// detector (the reference's blank_intensity, scan_simulator.cuh:34, 79).
template <class T>
void blank_intensity(em_plan p, const multislice_params<T>& m, std::vector<double>& per_det, double* total) {
  const size_t N = (size_t)m.nx * m.ny;
  Dev psi(N * 8);
  Dev out(sizeof(float) * (m.detector.size() + 1));
  const em_lens cl = lens_of(m.cond_lens);
  const double xb = 0, yb = 0;
  ok(em_build_probes(p, &cl, 0.0, &xb, &yb, 1, psi.p), p, "em_build_probes");
  ok(em_fft2d(p, psi.p, 1, -1), p, "em_fft2d");
  std::vector<double> in = det_angles(m.detector.ang_inner), ou = det_angles(m.detector.ang_outer);
  in.push_back(0.0);
  ou.push_back(0.0);  // outer <= 0: every frequency
  const int nd = (int)in.size();
  ok(em_detector_sums(p, psi.p, 1, nd, in.data(), ou.data(), 1.f, 0, out.as<float>()), p, "em_detector_sums");
  std::vector<float> h(nd);
  hip_ok(hipMemcpy(h.data(), out.p, sizeof(float) * nd, hipMemcpyDeviceToHost), "copy");
  per_det.assign(h.begin(), h.end() - 1);
  *total = h.back();
}

}  // namespace

template <class T>
void multislice_params<T>::validate_parameters() const {
  if (!em_fft_size_supported(nx) || !em_fft_size_supported(ny)) throw simulation_error("grid size not supported");
  if (!(lx > 0) || !(ly > 0) || !(dz > 0) || !(E_0 > 0)) throw simulation_error("grid or beam parameters");
  if (atoms.size() == 0) throw simulation_error("no atoms");
  if (detector.ang_inner.size() != detector.ang_outer.size() || detector.size() > 8)
    throw simulation_error("detectors: matching inner/outer angles, at most 8");
  if (simulation_type == eTEMST_STEM && (STEM_col <= 0 || STEM_row <= 0 || detector.size() == 0))
    throw simulation_error("STEM needs a scan and detectors");
  if (frozen_phonon_enable() && pn_nconf <= 0) throw simulation_error("pn_nconf");
  if (illumination_model != eIM_Coherent && illumination_model != eIM_Partial_Coherent &&
      illumination_model != eIM_Trans_Cross_Coef)
    throw simulation_error("illumination model");
}

template <class T>
T multislice_params<T>::output_lamda() const {
  double lam = 0, gam = 0, sig = 0;
  ok(em_host_constants((double)E_0, &lam, &gam, &sig), nullptr, "em_host_constants");
  return (T)lam;
}
template <class T>
T multislice_params<T>::output_gama() const {
  double lam = 0, gam = 0, sig = 0;
  ok(em_host_constants((double)E_0, &lam, &gam, &sig), nullptr, "em_host_constants");
  return (T)gam;
}
template <class T>
T multislice_params<T>::output_sigma() const {
  double lam = 0, gam = 0, sig = 0;
  ok(em_host_constants((double)E_0, &lam, &gam, &sig), nullptr, "em_host_constants");
  return (T)sig;
}

template <class T>
void vibrate_atom_positions(const Atom_Coord_Data<T>& in, Atom_Coord_Data<T>& out, unsigned long long seed,
                            int iconf, T sigma, const bool dims[3]) {
  const uint64_t g = 0x9E3779B97F4A7C15ull;
  uint64_t state = seed * g + (uint64_t)iconf;
  auto next = [&]() {
    state += g;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const size_t n = in.size();
  out = in;
  double pending = 0;
  bool have = false;
  auto normal = [&]() {
    if (have) {
      have = false;
      return pending;
    }
    const double u1 = 1.0 - (double)(next() >> 11) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(next() >> 11) * (1.0 / 9007199254740992.0);
    const double r = std::sqrt(-2.0 * std::log(u1));
    pending = r * std::sin(2 * 3.141592653589793238462643383279502884 * u2);
    have = true;
    return r * std::cos(2 * 3.141592653589793238462643383279502884 * u2);
  };
  for (size_t i = 0; i < n; ++i) {
    const double gx = normal(), gy = normal(), gz = normal();
    if (dims[0]) out.x[i] = (T)(in.x[i] + sigma * gx);
    if (dims[1]) out.y[i] = (T)(in.y[i] + sigma * gy);
    if (dims[2]) out.z[i] = (T)(in.z[i] + sigma * gz);
  }
}

// ED_HRTEM / CBED_CBEI (single_simulator.cuh:10-206) for one device.
template <class T>
void run_simu(multislice_params<T>* mp, int /*GPU_ID*/, output_Multi_Data<T>* o) {
  const multislice_params<T>& m = *mp;
  m.validate_parameters();
  Plan plan(grid_of(m), beam_of(m), 1);
  configure_potential(plan.p, m);
  const size_t N = (size_t)m.nx * m.ny;
  const bool full = (m.data_record == eWave);
  const int border = (m.simulation_type == eTEMST_HRTEM) ? EM_BORDER_MAX : EM_BORDER_MIN;
  int nconf;
  float w;
  phonon_range(m, &nconf, &w);
  o->init(m.ny, m.nx, 0, full);
  o->interval_x = m.lx / m.nx;
  o->interval_y = m.ly / m.ny;
  // Frozen-phonon configurations run as the probes of one batch, each with its
  // own potential stack (em_multislice_forward_multi), in blocks of at most
  // kMaxConfBlock configurations and about kConfBlockBytes of potentials; the
  // reference runs them one after another (single_simulator.cuh:60-80).
  constexpr int kMaxConfBlock = 64;
  constexpr size_t kConfBlockBytes = size_t(8) << 30;
  const int kmax = std::min(nconf, kMaxConfBlock);
  Dev psi(N * 8 * (size_t)kmax), a(N * 4), b(N * 4);
  hip_ok(hipMemset(a.p, 0, N * 4), "memset");
  hip_ok(hipMemset(b.p, 0, N * 4), "memset");
  const em_lens obj = lens_of(m.obj_lens), cond = lens_of(m.cond_lens);
  for (int iconf = 1; iconf <= nconf;) {
    std::vector<std::unique_ptr<Dev>> V;
    std::vector<const float*> Vp;
    std::vector<int> Ss;
    size_t used = 0;
    while (iconf <= nconf && (int)V.size() < kmax && used < kConfBlockBytes) {
      const Atom_Coord_Data<T> at = atoms_of_config(m, iconf);
      const int S = set_specimen(plan.p, at, border);
      V.push_back(std::make_unique<Dev>((size_t)S * N * 4));
      ok(em_potential_slices(plan.p, 0, S, V.back()->as<float>()), plan.p, "em_potential_slices");
      Vp.push_back(V.back()->as<float>());
      Ss.push_back(S);
      used += (size_t)S * N * 4;
      void* slot = psi.as<float2>() + (long long)(V.size() - 1) * (long long)N;
      if (m.iw_type == eIWT_Convergent_Wave) {
        // the probe's defocus is referred to this configuration's first slice
        const double xb = m.incident_wave_x_pos, yb = m.incident_wave_y_pos;
        ok(em_build_probes(plan.p, &cond, zmid0(plan.p), &xb, &yb, 1, slot), plan.p, "em_build_probes");
      } else {
        ok(em_plane_wave(plan.p, 1, slot), plan.p, "em_plane_wave");
      }
      ++iconf;
    }
    const int k = (int)V.size();
    if (k == 1)
      ok(em_multislice_forward(plan.p, Vp[0], nullptr, Ss[0], psi.p, 1, nullptr, 1), plan.p, "em_multislice_forward");
    else
      ok(em_multislice_forward_multi(plan.p, Vp.data(), Ss.data(), psi.p, k), plan.p, "em_multislice_forward_multi");
    for (int j = 0; j < k; ++j) {
      void* wj = psi.as<float2>() + (long long)j * (long long)N;
      if (m.simulation_type == eTEMST_HRTEM) {
        // Lens_Modulation::wave_modulation's illumination switch (lens_modulation.cuh:52-73)
        float* I = full ? nullptr : a.as<float>();
        float* re = full ? a.as<float>() : nullptr;
        float* im = full ? b.as<float>() : nullptr;
        em_coherence coh;
        coh.dsf_sigma = m.obj_lens.dsf_sigma;
        coh.dsf_npoints = m.obj_lens.dsf_npoints;
        coh.ssf_beta = m.obj_lens.ssf_beta;
        coh.ssf_npoints = m.obj_lens.ssf_npoints;
        coh.mode = (int)m.temporal_spatial_incoh;
        if (m.illumination_model == eIM_Partial_Coherent)
          ok(em_pctf_image(plan.p, &obj, &coh, wj, 1, I, re, im, w, 1), plan.p, "em_pctf_image");
        else if (m.illumination_model == eIM_Trans_Cross_Coef)
          ok(em_tcc_image(plan.p, &obj, &coh, wj, 1, I, re, im, w, 1), plan.p, "em_tcc_image");
        else
          ok(em_ctf_image(plan.p, &obj, wj, 1, I, re, im, w, 1), plan.p, "em_ctf_image");
      } else if (full) {
        hip_ok(em::launch_split(static_cast<const float2*>(wj), a.as<float>(), b.as<float>(), (long long)N, w, 1,
                                nullptr),
               "split");
      } else {
        hip_ok(em::launch_abs2(static_cast<const float2*>(wj), a.as<float>(), (long long)N, w, 1, nullptr), "abs2");
      }
    }
    hip_ok(hipDeviceSynchronize(), "sync");
  }
  if (m.simulation_type != eTEMST_HRTEM) {  // fft_shift of the diffraction planes
    Dev t(N * 4);
    hip_ok(em::launch_fft_shift(a.as<float>(), t.as<float>(), m.nx, m.ny, 1, 1.f, 0, nullptr), "shift");
    hip_ok(hipMemcpy(a.p, t.p, N * 4, hipMemcpyDeviceToDevice), "copy");
    if (full) {
      hip_ok(em::launch_fft_shift(b.as<float>(), t.as<float>(), m.nx, m.ny, 1, 1.f, 0, nullptr), "shift");
      hip_ok(hipMemcpy(b.p, t.p, N * 4, hipMemcpyDeviceToDevice), "copy");
    }
  }
  std::vector<float> ha(N), hb(N);
  hip_ok(hipMemcpy(ha.data(), a.p, N * 4, hipMemcpyDeviceToHost), "copy");
  if (full) {
    hip_ok(hipMemcpy(hb.data(), b.p, N * 4, hipMemcpyDeviceToHost), "copy");
    o->image_real.assign(ha.begin(), ha.end());
    o->image_imag.assign(hb.begin(), hb.end());
  } else {
    o->image_intensity.assign(ha.begin(), ha.end());
  }
}

// run_STEM_tile_simu / run_STEM_general_simu (scan_simulator.cuh:10-183).
template <class T>
void run_STEM_simu(multislice_params<T>* mp, int /*GPU_ID*/, output_Multi_Data<T>* o) {
  const multislice_params<T>& m = *mp;
  m.validate_parameters();
  const int nd = (int)m.detector.size();
  std::vector<double> x, y;
  scan_positions(m, x, y);
  const int npos = (int)x.size();
  const size_t N = (size_t)m.nx * m.ny;
  int nconf;
  float w;
  phonon_range(m, &nconf, &w);
  const bool tile = (m.stem_scanning == eTile_scanning);
  Plan plan(grid_of(m), beam_of(m), tile ? std::max(1, m.probe_batch) : 1);
  configure_potential(plan.p, m);
  o->init(m.STEM_row, m.STEM_col, nd, false);
  o->interval_x = (m.scan_range[2] - m.scan_range[0]) * m.lx / m.STEM_col;
  o->interval_y = (m.scan_range[3] - m.scan_range[1]) * m.ly / m.STEM_row;
  std::vector<double> blank;
  double total = 0;
  blank_intensity(plan.p, m, blank, &total);
  o->total_intensity = (T)total;
  const std::vector<double> in = det_angles(m.detector.ang_inner), ou = det_angles(m.detector.ang_outer);
  const em_lens cond = lens_of(m.cond_lens);
  Dev out(sizeof(float) * (size_t)nd * npos);
  hip_ok(hipMemset(out.p, 0, sizeof(float) * (size_t)nd * npos), "memset");
  std::vector<float> host((size_t)nd * npos, 0.f);
  if (tile) {
    for (int iconf = 1; iconf <= nconf; ++iconf) {
      const Atom_Coord_Data<T> at = atoms_of_config(m, iconf);
      set_specimen(plan.p, at, EM_BORDER_MIN);
      ok(em_stem(plan.p, &cond, npos, x.data(), y.data(), nd, in.data(), ou.data(), w, 1, out.as<float>()), plan.p,
         "em_stem");
    }
    hip_ok(hipMemcpy(host.data(), out.p, sizeof(float) * host.size(), hipMemcpyDeviceToHost), "copy");
  } else {
    // general scanning: the atoms are re-cut around every probe and the
    // probe is built at the window centre (multislice_params_dev.cuh:159).
    // Each probe keeps its own potential stack and defocus; blocks of probes
    // then run through their slices as the waves of one batch
    // (em_multislice_forward_multi) instead of one wave per launch.
    constexpr int kMaxProbeBlock = 16;
    constexpr size_t kBlockBytes = size_t(8) << 30;
    Dev psi(N * 8 * kMaxProbeBlock), d(sizeof(float) * nd * kMaxProbeBlock);
    std::vector<float> hd((size_t)nd * kMaxProbeBlock);
    for (int iconf = 1; iconf <= nconf; ++iconf) {
      const Atom_Coord_Data<T> at = atoms_of_config(m, iconf);
      for (int ip = 0; ip < npos;) {
        std::vector<std::unique_ptr<Dev>> V;
        std::vector<const float*> Vp;
        std::vector<int> Ss, pos;
        size_t used = 0;
        while (ip < npos && (int)pos.size() < kMaxProbeBlock && used < kBlockBytes) {
          const double xs = x[ip] - m.lx / 2, ys = y[ip] - m.ly / 2;
          Atom_Coord_Data<T> cut;
          for (size_t i = 0; i < at.size(); ++i)
            if (at.x[i] >= xs - 2.0 && at.x[i] <= xs + m.lx + 2.0 && at.y[i] >= ys - 2.0 &&
                at.y[i] <= ys + m.ly + 2.0)
              cut.push_back(at.Z[i], (T)(at.x[i] - xs), (T)(at.y[i] - ys), at.z[i], at.occ.empty() ? T(1) : at.occ[i]);
          if (cut.size() == 0) {
            for (int k = 0; k < nd; ++k) host[(size_t)k * npos + ip] += w * (float)blank[k];
            ++ip;
            continue;
          }
          const int S = set_specimen(plan.p, cut, EM_BORDER_MIN);
          V.emplace_back(new Dev((size_t)S * N * 4));
          ok(em_potential_slices(plan.p, 0, S, V.back()->as<float>()), plan.p, "em_potential_slices");
          const double xb = 0, yb = 0;
          ok(em_build_probes(plan.p, &cond, zmid0(plan.p), &xb, &yb, 1, psi.as<float>() + 2 * N * pos.size()),
             plan.p, "em_build_probes");
          Vp.push_back(V.back()->as<float>());
          Ss.push_back(S);
          pos.push_back(ip);
          used += (size_t)S * N * 4;
          ++ip;
        }
        const int k = (int)pos.size();
        if (k == 0) continue;
        ok(em_multislice_forward_multi(plan.p, Vp.data(), Ss.data(), psi.p, k), plan.p,
           "em_multislice_forward_multi");
        ok(em_detector_sums(plan.p, psi.p, k, nd, in.data(), ou.data(), w, 0, d.as<float>()), plan.p,
           "em_detector_sums");
        hip_ok(hipMemcpy(hd.data(), d.p, sizeof(float) * nd * k, hipMemcpyDeviceToHost), "copy");
        for (int b = 0; b < k; ++b)
          for (int kk = 0; kk < nd; ++kk) host[(size_t)kk * npos + pos[b]] += hd[(size_t)kk * k + b];
      }
    }
  }
  o->STEM_intensity.assign(host.begin(), host.end());
}

// run_thickness_simu (thickness_simulator.cuh:11-85), full-wave copy.
template <class T>
void run_thickness_simu(multislice_params<T>* mp, int /*GPU_ID*/, output_Multi_Data<T>* o) {
  const multislice_params<T>& m = *mp;
  m.validate_parameters();
  if ((int)m.detector.size() <= m.thickness_det) throw simulation_error("thickness filter detector missing");
  Plan plan(grid_of(m), beam_of(m), 1);
  configure_potential(plan.p, m);
  const size_t N = (size_t)m.nx * m.ny;
  int nconf;
  float w;
  phonon_range(m, &nconf, &w);
  o->row = m.ny;
  o->col = m.nx;
  const em_lens cond = lens_of(m.cond_lens);
  // Frozen-phonon configurations with one slice count run as the waves of one
  // batch (em_thickness_multi), in blocks as in run_simu; their frames are
  // summed into the output in configuration order.
  constexpr int kMaxConfBlock = 64;
  constexpr size_t kConfBlockBytes = size_t(8) << 30;
  const int kmax = std::min(nconf, kMaxConfBlock);
  Dev psi(N * 8 * (size_t)kmax);
  std::unique_ptr<Dev> frames, scratch;
  size_t scratch_bytes = 0;
  int nf_max = 0, nf = 0;
  const int st = m.thickness_stride;
  auto frames_of = [st](int S) { return st <= 1 ? S : S / st + (S % st != 0 ? 1 : 0); };
  for (int iconf = 1; iconf <= nconf;) {
    std::vector<std::unique_ptr<Dev>> V;
    std::vector<const float*> Vp;
    size_t used = 0;
    int S0 = 0;
    while (iconf <= nconf && (int)V.size() < kmax && used < kConfBlockBytes) {
      const Atom_Coord_Data<T> at = atoms_of_config(m, iconf);
      const int S = set_specimen(plan.p, at, EM_BORDER_MIN);
      if (!frames) {
        nf_max = S;
        frames.reset(new Dev(sizeof(float) * N * nf_max));
        hip_ok(hipMemset(frames->p, 0, sizeof(float) * N * nf_max), "memset");
      }
      if (S > nf_max) throw simulation_error("slice count changed between phonon configurations");
      if (!V.empty() && S != S0) break;  // this configuration opens the next block
      S0 = S;
      V.push_back(std::make_unique<Dev>((size_t)S * N * 4));
      ok(em_potential_slices(plan.p, 0, S, V.back()->as<float>()), plan.p, "em_potential_slices");
      Vp.push_back(V.back()->as<float>());
      used += ((size_t)S + (size_t)frames_of(S)) * N * 4;  // potentials and this configuration's frames
      void* slot = psi.as<float2>() + (long long)(V.size() - 1) * (long long)N;
      if (m.iw_type == eIWT_Convergent_Wave) {
        const double xb = m.incident_wave_x_pos, yb = m.incident_wave_y_pos;
        ok(em_build_probes(plan.p, &cond, zmid0(plan.p), &xb, &yb, 1, slot), plan.p, "em_build_probes");
      } else {
        ok(em_plane_wave(plan.p, 1, slot), plan.p, "em_plane_wave");
      }
      ++iconf;
    }
    const int k = (int)V.size();
    const double inner = m.detector.ang_inner[m.thickness_det], outer = m.detector.ang_outer[m.thickness_det];
    if (k == 1) {
      ok(em_thickness(plan.p, Vp[0], S0, psi.p, 1, inner, outer, m.thickness_stride, w, 1, frames->as<float>(), &nf),
         plan.p, "em_thickness");
    } else {
      const size_t need = sizeof(float) * N * (size_t)frames_of(S0) * (size_t)k;
      if (need > scratch_bytes) {
        scratch.reset();
        scratch.reset(new Dev(need));
        scratch_bytes = need;
      }
      ok(em_thickness_multi(plan.p, Vp.data(), S0, psi.p, k, inner, outer, m.thickness_stride, w, 0,
                            scratch->as<float>(), &nf),
         plan.p, "em_thickness_multi");
      // frames[f] += sum over the block's configurations (a deterministic reduction)
      for (int f = 0; f < nf; ++f)
        hip_ok(em::launch_grad_reduce((long long)N / 2, k,
                                      reinterpret_cast<const float2*>(scratch->as<float>() + (size_t)f * k * N),
                                      reinterpret_cast<float2*>(frames->as<float>() + (size_t)f * N), 1, nullptr),
               "frame sum");
    }
    hip_ok(hipDeviceSynchronize(), "sync");
  }
  std::vector<float> h(N * nf);
  hip_ok(hipMemcpy(h.data(), frames->p, sizeof(float) * h.size(), hipMemcpyDeviceToHost), "copy");
  o->thickness_frames.assign(h.begin(), h.end());
}

template <class T>
void TEM_simulator<T>::TEM_multiGPU() {
  run_simu<T>(multisliceparam_host_ptr, 0, output_mulData_ptr);
}
template <class T>
void scan_simulator<T>::STEM_multiGPU() {
  run_STEM_simu<T>(multisliceparam_host_ptr, 0, output_mulData_ptr);
}
template <class T>
void Thickness_simulator<T>::Thickness_multiGPU() {
  run_thickness_simu<T>(multisliceparam_host_ptr, 0, output_mulData_ptr);
}

#define EM_INSTANTIATE(T)                                                                                     \
  template struct multislice_params<T>;                                                                       \
  template class TEM_simulator<T>;                                                                            \
  template class scan_simulator<T>;                                                                           \
  template class Thickness_simulator<T>;                                                                      \
  template void run_simu<T>(multislice_params<T>*, int, output_Multi_Data<T>*);                               \
  template void run_STEM_simu<T>(multislice_params<T>*, int, output_Multi_Data<T>*);                          \
  template void run_thickness_simu<T>(multislice_params<T>*, int, output_Multi_Data<T>*);                     \
  template void vibrate_atom_positions<T>(const Atom_Coord_Data<T>&, Atom_Coord_Data<T>&, unsigned long long, \
                                          int, T, const bool*);
EM_INSTANTIATE(float)
EM_INSTANTIATE(double)
#undef EM_INSTANTIATE

}  // namespace cudaEM

// mulSimulator (simulator.cu:7-49)
void mulSimulator::executionTEM(const cudaEM::multislice_params<float>& in, cudaEM::output_Multi_Data<float>* out) {
  cudaEM::multislice_params<float> p;
  p.assign(in);
  p.validate_parameters();
  p.initial_OL();
  cudaEM::TEM_simulator<float> s;
  s.set_input_data(&p, out);
  s.TEM_multiGPU();
}

void mulSimulator::executionCBED(const cudaEM::multislice_params<float>& in, cudaEM::output_Multi_Data<float>* out) {
  cudaEM::multislice_params<float> p;
  p.assign(in);
  p.validate_parameters();
  p.initial_CL();
  cudaEM::TEM_simulator<float> s;
  s.set_input_data(&p, out);
  s.TEM_multiGPU();
}

void mulSimulator::executionSTEM(const cudaEM::multislice_params<float>& in, cudaEM::output_Multi_Data<float>* out) {
  cudaEM::multislice_params<float> p;
  p.assign(in);
  p.is_crystal = false;
  p.simulation_type = cudaEM::eTEMST_STEM;
  p.validate_parameters();
  p.initial_CL();
  p.set_Scanning({0.0f, 0.0f, 1.0f, 1.0f});
  cudaEM::scan_simulator<float> s;
  s.set_input_data(&p, out);
  s.STEM_multiGPU();
}

void mulSimulator::executionThickness(const cudaEM::multislice_params<float>& in,
                                      cudaEM::output_Multi_Data<float>* out) {
  cudaEM::multislice_params<float> p;
  p.assign(in);
  p.validate_parameters();
  cudaEM::Thickness_simulator<float> s;
  s.set_input_data(&p, out);
  s.Thickness_multiGPU();
}
