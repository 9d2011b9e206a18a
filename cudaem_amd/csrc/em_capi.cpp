// em_capi.cpp -- the C ABI (include/cudaem_amd.h): plans, specimen upload,
// operator entry points and the fused multislice pipelines.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "cudaem_amd.h"
#include "em_host.hpp"
#include "em_kernels.hpp"

using namespace em;

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kZEps = 1e-4;  // epsilon_rel of the slicer (float eps of misc_dev.cuh:17-18)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    if (p) {
      // work of an earlier call may still read the old buffer (em_stem
      // returns with its graphs in flight)
      (void)hipDeviceSynchronize();
      (void)hipFree(p);
      p = nullptr;
      bytes = 0;
    }
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

struct Graph {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  ~Graph() {
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
  }
};

// Everything a captured STEM batch bakes into its launches: the batch shape,
// detector bands, launch shapes, band skip and the device buffers it reads
// and writes.  Zero-filled before use, so it compares bytewise.
struct StemGraphKey {
  int nb, S, ndet, skip_lo, skip_hi, skip_lo_y, skip_hi_y;
  em_tuning tune;
  double dg2min[kMaxDet], dg2max[kMaxDet];
  const void* buf[10];
  bool operator==(const StemGraphKey& o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

// Pinned host memory (hipHostMalloc), for asynchronous uploads.
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    if (p) {
      (void)hipHostFree(p);
      p = nullptr;
      bytes = 0;
    }
    hipError_t e = hipHostMalloc(&p, b, hipHostMallocDefault);
    if (e == hipSuccess) bytes = b;
    return e;
  }
};

// Events owned by one call: destroyed on every return path.
struct EventSet {
  std::vector<hipEvent_t> ev;
  EventSet() = default;
  EventSet(const EventSet&) = delete;
  EventSet& operator=(const EventSet&) = delete;
  ~EventSet() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
  hipError_t make(hipEvent_t* out, unsigned flags = hipEventDefault) {
    hipEvent_t e = nullptr;
    const hipError_t r = hipEventCreateWithFlags(&e, flags);
    if (r == hipSuccess) {
      ev.push_back(e);
      *out = e;
    }
    return r;
  }
};

// Fixed device scratch of the allocation-free entry points: the detector
// pass's per-tile partial sums, [nprobe][ndet][ntiles] floats.  Each device
// holds its own copy until the process ends; host code uses the address
// hipGetSymbolAddress reports for the current device, never the name.
constexpr size_t kNoAllocScratchFloats = size_t(1) << 18;  // 1 MiB
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};

// The current device's scratch address, looked up once per device.
hipError_t device_scratch(int dev, float** out) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<float*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace

struct em_plan_s {
  em_grid grid{};
  em_beam beam{};
  int max_batch = 1;
  int device = 0;
  float* scratch = nullptr;  // this device's fixed scratch (allocation-free entry points)
  hipStream_t stream = nullptr;
  Constants k{};
  double vr_factor = 0, gx0 = 0, gy0 = 0;
  GridDev gdev{};
  em_tuning tune{4096, 8192, 0, 0, 0, 0, 0, 0, 0};
  // band-limit skipping: columns kx in [skip_lo, skip_hi] of the propagator
  // are below band_skip (relative) for every ky; skip_lo > skip_hi: none
  double band_skip = 1e-9;
  int skip_lo = 1, skip_hi = 0;
  // the same range along ky (rows of a column whose propagator is below
  // band_skip whatever kx is), and the line slots the pruned passes may drop
  int skip_lo_y = 1, skip_hi_y = 0;
  int zk_x = 0, zk_y = 0;
  std::string err;
  // spline settings and element overrides
  int nR = 128;
  double r_min = 0.005, r_max = 3.0, tap = 0.85;
  std::map<int, ElementCoefs> custom;
  // projected-potential splines by atomic number, for the current
  // coefficients and spline parameters (a specimen change rebuilds none)
  std::map<int, Spline> spline_cache;
  // fine slicing (projected_potential.cuh:73-91): a spline per atom entry of
  // every slice, from the element's 3D potential integrated over the slice
  struct Params3D {
    int pt;
    double cl[6], cnl[6];
  };
  bool fine = false;
  int nqz = 128;
  std::map<int, Params3D> params3d;
  long long fine_entries = 0;  // atom entries of the current specimen's splines
  // specimen (host)
  bool has_specimen = false;
  std::vector<double> ax, ay, az, aocc;
  std::vector<int> aZ;
  std::vector<int> types;  // distinct Z, ascending
  std::vector<Spline> splines;
  std::vector<SliceRange> slices;
  int nbx = 0, nby = 0;
  float bin = 0;
  // device
  DevBuf tw_x, tw_y, P, atoms, off, tables, rmax2;
  DevBuf twt_x, twt_y;  // per-lane twiddle tables of radix variant 3 (1024 only)
  DevBuf wave, tmp, dpart, A, ex, ey, xe, ye, gpart, Vstore, Tstore, detout, seg, segwave;
  DevBuf fine_tab, fine_az, fine_el, fine_r2, fine_q;
  DevBuf pk_ctr, pk_err;  // persistent slice loop: work queues and barrier count, error flag
  int pk_grid = 0;        // its workgroups (0: not yet queried)
  DevBuf mchunks;  // chunk tables of em_multislice_forward_multi
  int vstore_slices = -1;
  hipStream_t gstream = nullptr;  // non-blocking stream for graph capture
  // em_stem: instantiated batch graphs kept across calls (a phonon step
  // re-launches them), the probe positions of a call staged in pinned memory
  // and uploaded once, and the events that order the graph stream after the
  // caller's stream and back
  std::vector<std::pair<StemGraphKey, std::unique_ptr<Graph>>> stem_graphs;
  PinnedBuf pos_host;
  DevBuf pos_dev;
  hipEvent_t pos_copied = nullptr, stem_in = nullptr, stem_out = nullptr;
  // em_multislice_ragged (on the plan of request 0): instantiated graphs of
  // request sets kept across calls, each with its chunk table, keyed by
  // everything a capture bakes in; and the events that order the graph
  // stream after the caller's stream and back
  struct RaggedGraph {
    std::string key;
    Graph graph;
    DevBuf chunks;
  };
  std::vector<std::unique_ptr<RaggedGraph>> ragged_graphs;
  hipEvent_t rag_in = nullptr, rag_out = nullptr;
  // per-launch timing of the STEM slice loop's row / column passes: event
  // pairs (start, end) of each timed launch, kind 0 = row, 1 = column
  bool pass_timing = false;
  std::vector<hipEvent_t> tev;
  std::vector<int> tkind;
  void clear_pass_events() {
    for (hipEvent_t e : tev) (void)hipEventDestroy(e);
    tev.clear();
    tkind.clear();
  }
  ~em_plan_s() {
    // a ragged graph kept on another plan may still read this plan's tables
    (void)hipDeviceSynchronize();
    clear_pass_events();
    stem_graphs.clear();
    ragged_graphs.clear();
    for (hipEvent_t e : {pos_copied, stem_in, stem_out, rag_in, rag_out})
      if (e) (void)hipEventDestroy(e);
    if (gstream) (void)hipStreamDestroy(gstream);
  }

  em_status fail(em_status s, const std::string& m) {
    err = m;
    return s;
  }
  em_status hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return EM_OK;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return EM_ERR_HIP;
  }
  long long N() const { return (long long)grid.nx * grid.ny; }
};

#define EM_TRY(expr)                 \
  do {                               \
    em_status _s = (expr);           \
    if (_s != EM_OK) return _s;      \
  } while (0)
#define EM_HIP(plan, expr) EM_TRY((plan)->hip((expr), #expr))

namespace {

em_status upload_twiddles(em_plan p, DevBuf& buf, int n) {
  std::vector<float2> h(n);
  for (int m = 0; m < n; ++m) {
    const double a = -2.0 * kPi * (double)m / (double)n;
    h[m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  EM_HIP(p, buf.ensure(sizeof(float2) * n));
  EM_HIP(p, hipMemcpy(buf.p, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  return EM_OK;
}

// Per-lane twiddle tables of radix variant 3 at 1024 (the kernels' TanRow1024
// and TanCol1024 describe the layouts), built in FP64 and rounded once.  A
// stage twiddle exp(DIR i theta) is stored as T = tan(theta) and g = cos(theta),
// plus the ratios of g's that ride on the radix-4 FMAs.
em_status upload_tangent_tables(em_plan p, DevBuf& buf, int n, bool rows) {
  if (n != 1024) return EM_OK;
  std::vector<float> h;
  // radix 4 of stage size rns: T1, T2, T3, g1, g2, g3/g1, 0, 0 for class k
  auto radix4 = [&h](int k, int step, int rns) {
    double T[4], c[4];
    for (int r = 1; r <= 3; ++r) {
      const double th = 2.0 * kPi * (double)r * step * k / rns;
      T[r] = std::tan(th);
      c[r] = std::cos(th);
    }
    for (double x : {T[1], T[2], T[3], c[1], c[2], c[3] / c[1]}) h.push_back((float)x);
  };
  if (rows) {
    // stage 2 (radix 16 over 256): 16 classes of 40 floats
    auto G16 = [](int m) { return (m % 16) % 4 == 0 ? 1.0 : std::cos(2.0 * kPi * (m % 16) / 16.0); };
    for (int k = 0; k < 16; ++k) {
      double c[16], T[16];
      for (int r = 0; r < 16; ++r) {
        const double th = 2.0 * kPi * r * k / 256.0;
        c[r] = std::cos(th);
        T[r] = std::tan(th);
      }
      for (int r = 1; r < 16; ++r) h.push_back((float)T[r]);
      h.push_back(0.f);
      for (int n2 = 0; n2 < 4; ++n2) h.push_back((float)(c[n2 + 8] / c[n2]));
      for (int n2 = 0; n2 < 4; ++n2) h.push_back((float)(c[n2 + 12] / c[n2 + 4]));
      for (int n2 = 0; n2 < 4; ++n2) h.push_back((float)(c[n2 + 4] / c[n2]));
      for (int k1 = 0; k1 < 4; ++k1) h.push_back((float)(c[2] * G16(2 * k1)));
      for (int k1 = 0; k1 < 4; ++k1) h.push_back((float)((c[3] * G16(3 * k1)) / (c[1] * G16(k1))));
      for (int k1 = 0; k1 < 4; ++k1) h.push_back((float)(c[1] * G16(k1)));
    }
    // stage 3 (radix 4 over 1024): 256 classes
    for (int k = 0; k < 256; ++k) {
      radix4(k, 1, 1024);
      h.push_back(0.f);
      h.push_back(0.f);
    }
  } else {
    // stage 2 (radix 4 over 64): 16 classes
    for (int k = 0; k < 16; ++k) {
      radix4(k, 1, 64);
      h.push_back(0.f);
      h.push_back(0.f);
    }
    // stage 3 (radix 16 over 1024): 64 classes of lo_1..lo_3 and W^(4k)'s radix 4
    for (int k = 0; k < 64; ++k) {
      for (int m = 1; m <= 3; ++m) {
        const double a = -2.0 * kPi * m * k / 1024.0;
        h.push_back((float)std::cos(a));
        h.push_back((float)std::sin(a));
      }
      radix4(k, 4, 1024);
    }
  }
  EM_HIP(p, buf.ensure(sizeof(float) * h.size()));
  EM_HIP(p, hipMemcpy(buf.p, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
  return EM_OK;
}

// Probes per workgroup: enough workgroups to fill 256 CUs several times over,
// and as many probes as possible sharing one register-held t or P.
int row_pb(const em_plan p, int B) {
  const int blocks = p->grid.ny / row_lines_per_block(p->grid.nx);
  const int target = std::max(1, p->tune.row_blocks);
  int groups = std::max(1, std::min(B, (target + blocks - 1) / blocks));
  return (B + groups - 1) / groups;
}
int col_pb(const em_plan p, int B) {
  const int tile = p->tune.col_tile ? p->tune.col_tile : col_tile(p->grid.ny);
  const int blocks = p->grid.nx / tile;
  const int target = std::max(1, p->tune.col_blocks);
  int groups = std::max(1, std::min(B, (target + blocks - 1) / blocks));
  return (B + groups - 1) / groups;
}

// Columns whose band-width-limit factor 1/(1+exp(alpha (g^2 - gl^2))) is at
// most band_skip for every ky (the factor is largest at ky = 0): the
// propagator there is below band_skip / N, and a wave component passing it
// changes the result by less than band_skip relative -- under FP32 resolution
// for band_skip <= 1e-9.  They form the range |kx_s| >= K of the unshifted
// index, i.e. kx in [K, nx - K].
// The same along ky: the factor is largest at kx = 0, so rows |ky_s| >= K_y
// of every column are below band_skip.  The pruned passes drop the line
// slots wholly inside these ranges.
void compute_band_skip(em_plan p) {
  p->skip_lo = p->skip_lo_y = 1;
  p->skip_hi = p->skip_hi_y = 0;
  p->zk_x = p->zk_y = 0;
  if (!p->grid.bwl || !(p->band_skip > 0) || std::fabs(p->grid.dz) <= 1e-5) return;
  auto range = [&](int n, double dg, int* lo, int* hi) {
    for (int k = 1; k <= n / 2; ++k) {
      const double g = k * dg;
      const double f = 1.0 / (1.0 + std::exp(p->gdev.alpha * (g * g - p->gdev.gl2_max)));
      if (f <= p->band_skip) {
        *lo = k;
        *hi = n - k;
        return;
      }
    }
  };
  range(p->grid.nx, p->gdev.dgx, &p->skip_lo, &p->skip_hi);
  range(p->grid.ny, p->gdev.dgy, &p->skip_lo_y, &p->skip_hi_y);
  p->zk_x = pruned_slots(p->grid.nx, p->skip_lo, p->skip_hi);
  p->zk_y = pruned_slots(p->grid.ny, p->skip_lo_y, p->skip_hi_y);
}

RowArgs row_args(em_plan p, const float2* in, float2* out, int B) {
  RowArgs a{};
  a.in = in;
  a.out = out;
  a.pstride = p->N();
  a.nx = p->grid.nx;
  a.ny = p->grid.ny;
  a.nprobe = B;
  a.pb = row_pb(p, B);
  a.tw = p->tw_x.as<float2>();
  a.twt = p->twt_x.as<float2>();
  a.wgt = 1.f;
  a.skip_lo = p->skip_lo;
  a.skip_hi = p->skip_hi;
  a.skip_in = 0;
  a.skip_out = 0;
  a.radix = p->tune.row_radix;
  return a;
}

// the row pass of a slice step: input from a propagator column pass (or the
// probe build) and / or output to one
RowArgs row_args_skip(em_plan p, const float2* in, float2* out, int B, bool skip_in, bool skip_out) {
  RowArgs a = row_args(p, in, out, B);
  a.skip_in = skip_in ? 1 : 0;
  a.skip_out = skip_out ? 1 : 0;
  a.zk = (skip_in && skip_out) ? p->zk_x : 0;
  return a;
}

ColArgs col_args(em_plan p, const float2* in, float2* out, int B) {
  ColArgs a{};
  a.in = in;
  a.out = out;
  a.pstride = p->N();
  a.nx = p->grid.nx;
  a.ny = p->grid.ny;
  a.nprobe = B;
  a.pb = col_pb(p, B);
  a.P = p->P.as<float2>();
  a.tw = p->tw_y.as<float2>();
  a.twt = p->twt_y.as<float2>();
  a.grid = p->gdev;
  a.cols = p->tune.col_tile;
  const bool tr = col_transposed_available(p->grid.ny) && (a.cols == 0 || a.cols == 16 || a.cols == 8);
  a.variant = (p->tune.col_variant == 2 && tr) ? COL_VARIANT_TRANSPOSED : COL_VARIANT_BLOCK;
  a.skip_lo = 1;
  a.skip_hi = 0;
  a.radix = p->tune.col_radix;
  return a;
}

// a propagator column pass whose band-limited columns are skipped
ColArgs col_args_skip(em_plan p, const float2* in, float2* out, int B) {
  ColArgs a = col_args(p, in, out, B);
  a.skip_lo = p->skip_lo;
  a.skip_hi = p->skip_hi;
  a.zk = p->zk_y;
  return a;
}

em_status row(em_plan p, unsigned mode, const RowArgs& a) {
  return p->hip(launch_row(mode, a, p->stream), "row pass");
}
em_status col(em_plan p, unsigned mode, const ColArgs& a) {
  return p->hip(launch_col(mode, a, p->stream), "column pass");
}

// Launch through `run` (one row or column pass) with a pair of timing events
// that the launch itself records, when the plan records per-launch pass times
// (kind 0 = row pass, 1 = column pass).
template <typename F>
em_status timed(em_plan p, int kind, F&& run) {
  if (!p->pass_timing) return run();
  // timing-only events: no system-scope cache writeback between the passes,
  // as between the passes of a graph launch
  hipEvent_t e[2];
  EM_HIP(p, hipEventCreateWithFlags(&e[0], hipEventDisableSystemFence));
  if (hipEventCreateWithFlags(&e[1], hipEventDisableSystemFence) != hipSuccess) {
    (void)hipEventDestroy(e[0]);
    return p->fail(EM_ERR_HIP, "event create");
  }
  p->tev.push_back(e[0]);
  p->tev.push_back(e[1]);
  p->tkind.push_back(kind);
  set_pass_events(e[0], e[1]);
  const em_status st = run();
  const bool taken = !pass_events_pending();
  set_pass_events(nullptr, nullptr);
  EM_TRY(st);
  if (!taken) return p->fail(EM_ERR_HIP, "timed launch started no pass");
  return EM_OK;
}

em_status ensure_specimen_potential(em_plan p);

// the interior slice step on precomputed transmissions may read t per probe
// instead of holding it in registers (tuning row_variant 1)
unsigned row_t_mode(em_plan p, unsigned m) {
  return (m == (ROW_PRE_IFFT | ROW_MUL_T | ROW_POST_FFT) && p->tune.row_variant == 1) ? (m | ROW_T_STREAM) : m;
}

// Thickness series (run_thickness_simu, thickness_simulator.cuh:11-85): the
// slice loop with the band-passed intensity recorded after every `stride`-th
// slice and after the last one.  row_step(s) runs slice s's row pass on wv.
template <typename RowStep>
em_status thickness_loop(em_plan p, int S, float2* wv, int nprobe, double inner, double outer, int stride, float w,
                         int accumulate, float* frames, int* n_frames, RowStep&& row_step) {
  const long long N = p->N();
  EM_HIP(p, p->tmp.ensure(sizeof(float2) * N * nprobe));
  int nf = 0;
  for (int s = 0; s < S; ++s) {
    EM_TRY(row_step(s));
    EM_TRY(col(p, COL_PRE_FFT | COL_MUL_P | COL_POST_IFFT | COL_STORE, col_args(p, wv, wv, nprobe)));
    const bool rec = stride <= 1 || ((s + 1) % stride == 0) || s == S - 1;
    if (!rec) continue;
    ColArgs c = col_args(p, wv, p->tmp.as<float2>(), nprobe);
    c.fg2min = mrad_to_g2(inner, p->k.lambda, false);
    c.fg2max = mrad_to_g2(outer, p->k.lambda, true);
    // the wave here is in the (y, kx) layout, F_x psi / nx (the unnormalised
    // inverse of the next row pass restores psi), so the band-pass scale 1/N
    // of Filter_wave becomes nx/N = 1/ny
    c.fscale = (float)(1.0 / (double)p->grid.ny);
    EM_TRY(col(p, COL_PRE_FFT | COL_MUL_FILTER | COL_POST_IFFT | COL_STORE, c));
    RowArgs ro = row_args(p, p->tmp.as<float2>(), nullptr, nprobe);
    ro.o1 = frames + (long long)nf * nprobe * N;
    ro.wgt = w;
    EM_TRY(row(p, ROW_PRE_IFFT | ROW_OUT_ABS2 | (accumulate ? ROW_OUT_ACCUM : 0u), ro));
    ++nf;
  }
  // back to real space
  EM_TRY(row(p, ROW_PRE_IFFT, row_args(p, wv, wv, nprobe)));
  if (n_frames) *n_frames = nf;
  return EM_OK;
}

// Whether the interior of a slice loop on precomputed transmissions runs as
// one persistent launch (tuning persistent, and a grid and skip range its
// pruned items support).
bool use_persistent(em_plan p, bool on_t, int S) {
  return p->tune.persistent && on_t && S >= 2 && p->tune.row_variant == 0 &&
         persist_supported(p->grid.nx, p->grid.ny, p->zk_x, p->zk_y);
}

// The propagator passes of slices 0..S-2 and the row passes of slices
// 1..S-2 of nb probes at wave, in one persistent launch: the caller runs
// slice 0's row pass before and slice S-1's row pass and last column pass
// after.  Work items follow the launch-shape targets: row_blocks row items of
// 8 rows, col_blocks column items of 8 columns.
em_status persistent_interior(em_plan p, const float2* t, int S, float2* wave, int nb) {
  if (p->pk_grid <= 0) EM_HIP(p, persist_grid(p->grid.nx, &p->pk_grid));
  const int n = p->grid.nx;
  const int ctr = persist_counters(S);
  EM_HIP(p, p->pk_ctr.ensure(sizeof(float2) * (size_t)((ctr + 1) / 2)));
  EM_HIP(p, launch_fill(p->pk_ctr.as<float2>(), (ctr + 1) / 2, make_float2(0.f, 0.f), p->stream));
  PersistArgs a{};
  a.wave = wave;
  a.pstride = p->N();
  a.n = n;
  a.nprobe = nb;
  a.S = S;
  a.t = t;
  a.P = p->P.as<float2>();
  a.twx = p->tw_x.as<float2>();
  a.twy = p->tw_y.as<float2>();
  a.skip_lo = p->skip_lo;
  a.skip_hi = p->skip_hi;
  a.tile_lo = (p->skip_lo + 7) / 8;
  a.tile_hi = (p->skip_hi + 1) / 8 - 1;
  a.zk_x = p->zk_x;
  a.zk_y = p->zk_y;
  const int row_units = n / 8;
  const int kept = n / 8 - (a.tile_lo <= a.tile_hi ? a.tile_hi - a.tile_lo + 1 : 0);
  const int rch = std::max(1, std::min(nb, (p->tune.row_blocks + row_units - 1) / row_units));
  const int cch = std::max(1, std::min(nb, (p->tune.col_blocks + kept - 1) / kept));
  a.row_pb = (nb + rch - 1) / rch;
  a.col_pb = (nb + cch - 1) / cch;
  a.ctr = reinterpret_cast<int*>(p->pk_ctr.p);
  a.err = reinterpret_cast<int*>(p->pk_err.p);
  a.spin_limit = 50000000;  // 0.5 s of the 100 MHz wall clock
  return p->hip(launch_persist(a, p->pk_grid, p->stream), "persistent slice loop");
}

// Before a call's persistent launches: clear the error flag.
em_status persistent_begin(em_plan p) {
  EM_HIP(p, p->pk_err.ensure(sizeof(float2)));
  EM_HIP(p, launch_fill(p->pk_err.as<float2>(), 1, make_float2(0.f, 0.f), p->stream));
  return EM_OK;
}

// After a call's persistent launches: wait for them and report a barrier
// that gave up.
em_status persistent_check(em_plan p) {
  int err = 0;
  EM_HIP(p, hipStreamSynchronize(p->stream));
  EM_HIP(p, hipMemcpy(&err, p->pk_err.p, sizeof(int), hipMemcpyDeviceToHost));
  if (err != 0) return p->fail(EM_ERR_HIP, "persistent slice loop: a grid barrier wait timed out");
  return EM_OK;
}

// Fused forward multislice over nb probes at wave (real space in, Fourier out).
em_status forward_group(em_plan p, const float* V, const float2* t, int S, float2* wave, int nb,
                        float2* ckpt, long long ckpt_slice_stride, int every, bool in_mixed_layout) {
  const bool pk = !ckpt && use_persistent(p, t != nullptr && V == nullptr, S);
  for (int s = 0; s < S; ++s) {
    if (pk && s > 0 && s < S - 1) continue;  // in the persistent launch
    if (pk && s == 0) {
      RowArgs r = row_args_skip(p, wave, wave, nb, in_mixed_layout, true);
      r.t = t;
      EM_TRY(row(p, ROW_POST_FFT | ROW_MUL_T | (in_mixed_layout ? ROW_PRE_IFFT : 0u), r));
      EM_TRY(persistent_interior(p, t, S, wave, nb));
      continue;
    }
    RowArgs r = row_args_skip(p, wave, wave, nb, s > 0 || in_mixed_layout, s < S - 1);
    unsigned m = ROW_POST_FFT;
    if (s > 0 || in_mixed_layout) m |= ROW_PRE_IFFT;
    if (V) {
      m |= ROW_MUL_V;
      r.V = V + (long long)s * p->N();
    } else {
      m |= ROW_MUL_T;
      r.t = t + (long long)s * p->N();
    }
    if (ckpt && s % every == 0) {
      m |= ROW_CKPT_STORE;
      r.ckpt = ckpt + (long long)(s / every) * ckpt_slice_stride;
    }
    EM_TRY(row(p, row_t_mode(p, m), r));
    ColArgs c = col_args_skip(p, wave, wave, nb);
    EM_TRY(col(p, s < S - 1 ? (COL_PRE_FFT | COL_MUL_P | COL_POST_IFFT | COL_STORE) : (COL_PRE_FFT | COL_STORE), c));
  }
  return EM_OK;
}

// Row-pass workgroups the device holds at once (occupancy x CUs), 0 if the
// query fails.
int row_slots(em_plan p) {
  int dev = 0, cus = 0, occ = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      row_blocks_per_cu(p->grid.nx, row_t_mode(p, ROW_PRE_IFFT | ROW_MUL_T | ROW_POST_FFT), &occ) != hipSuccess)
    return 0;
  return std::max(0, occ) * std::max(0, cus);
}

// Default row-pass target: the fewest probe chunks whose workgroups fill whole
// rounds of the resident slots.  1024^2 on 256 CUs: 3 chunks of 256
// workgroups, one round of 768 (5% faster per probe-slice than 4096, whose
// last round is partly empty); 2048^2: 3 chunks of 1024, four rounds.
int whole_round_row_target(em_plan p) {
  const int slots = row_slots(p);
  const int per_chunk = p->grid.ny / row_lines_per_block(p->grid.nx);
  if (slots <= 0 || per_chunk <= 0) return 4096;
  return per_chunk * (slots / std::gcd(per_chunk, slots));
}

int group_size(em_plan p, int B) {
  int g = p->tune.probe_group;
  if (g <= 0) {
    // about 320 MiB of waves (8 B per pixel) per group: enough workgroups per
    // pass to fill the chip, while most of a group's traffic between the two
    // passes of a slice still hits the 256 MiB last-level cache (measured
    // best at 512^2 to 2048^2)
    const long long bytes = p->N() * 8;
    g = (int)std::max<long long>(1, (320ll << 20) / bytes);
    // waves over 16 MiB (2048^2 and up) gain nothing from the cache, and their
    // column pass holds a CU per workgroup: long launches pay more (2048^2:
    // groups of 64 ran 16% faster than groups of 10)
    if (bytes > (16ll << 20)) g = std::max(g, 64);
  }
  g = std::min(B, g);
  // even out the groups of the batch: no small, inefficient tail group
  const int ngroups = (B + g - 1) / g;
  return (B + ngroups - 1) / ngroups;
}

}  // namespace

extern "C" {

int em_fft_size_supported(int n) { return fft_size_supported(n) ? 1 : 0; }

em_status em_plan_create(const em_grid* grid, const em_beam* beam, int max_batch, em_plan* out) {
  if (!grid || !beam || !out || max_batch <= 0) return EM_ERR_INVALID;
  *out = nullptr;
  if (!fft_size_supported(grid->nx) || !fft_size_supported(grid->ny)) return EM_ERR_UNSUPPORTED;
  if (!(grid->lx > 0) || !(grid->ly > 0) || !(beam->e0_kv > 0)) return EM_ERR_INVALID;
  auto p = std::make_unique<em_plan_s>();
  p->grid = *grid;
  p->beam = *beam;
  p->max_batch = max_batch;
  p->k = beam_constants(beam->e0_kv);
  p->vr_factor = p->k.sigma / std::cos(beam->theta);
  p->gx0 = std::sin(beam->theta) * std::cos(beam->phi) / p->k.lambda;
  p->gy0 = std::sin(beam->theta) * std::sin(beam->phi) / p->k.lambda;
  p->gdev = make_grid_dev(*grid);
  EM_HIP(p.get(), hipGetDevice(&p->device));
  EM_HIP(p.get(), device_scratch(p->device, &p->scratch));
  EM_HIP(p.get(), prepare_col_kernels(grid->ny));
  EM_HIP(p.get(), hipStreamCreateWithFlags(&p->gstream, hipStreamNonBlocking));
  for (hipEvent_t* e : {&p->pos_copied, &p->stem_in, &p->stem_out, &p->rag_in, &p->rag_out})
    EM_HIP(p.get(), hipEventCreateWithFlags(e, hipEventDisableTiming));
  EM_TRY(upload_twiddles(p.get(), p->tw_x, grid->nx));
  EM_TRY(upload_twiddles(p.get(), p->tw_y, grid->ny));
  EM_TRY(upload_tangent_tables(p.get(), p->twt_x, grid->nx, true));
  EM_TRY(upload_tangent_tables(p.get(), p->twt_y, grid->ny, false));
  // Propagator, built once per plan (wave_function.cuh:42-44)
  em_plan raw = p.get();
  EM_HIP(raw, p->P.ensure(sizeof(float2) * p->N()));
  if (std::fabs(grid->dz) <= 1e-5) {
    EM_HIP(raw, launch_fill(p->P.as<float2>(), p->N(), make_float2(1.f, 0.f), nullptr));
  } else {
    const double w = -kPi * p->k.lambda * grid->dz;
    EM_HIP(raw, launch_propagator(p->gdev, w, p->gx0, p->gy0, grid->bwl, p->P.as<float2>(), nullptr));
  }
  EM_HIP(raw, hipDeviceSynchronize());
  compute_band_skip(raw);
  p->tune.row_blocks = whole_round_row_target(raw);
  // radix sequences measured fastest untuned (batched STEM, A/B of the
  // pairs): 16.4.16 columns and rows at 1024 (with the rows' buffer
  // access, 16.4.16 rows are 2% faster than 16.16.4; the columns' two
  // sequences tie), 16.8.16 rows at 2048 (1% faster row pass)
  p->tune.col_radix = grid->ny == 1024 ? 1 : 0;
  // two probes per column-pass workgroup at 1024 (P loaded once for both):
  // 3% faster column passes than one probe each (batched STEM, 256-probe
  // batches: 89.8 against 92.8 us per 37-probe launch, same box)
  if (grid->ny == 1024) p->tune.col_blocks = 4096;
  p->tune.row_radix = (grid->nx == 1024 || grid->nx == 2048) ? 1 : 0;
  *out = p.release();
  return EM_OK;
}

em_status em_plan_set_band_skip(em_plan plan, double threshold, int* skip_lo, int* skip_hi) {
  if (!plan || threshold < 0 || threshold > 1e-6) return EM_ERR_INVALID;
  plan->band_skip = threshold;
  compute_band_skip(plan);
  if (skip_lo) *skip_lo = plan->skip_lo;
  if (skip_hi) *skip_hi = plan->skip_hi;
  return EM_OK;
}

em_status em_plan_destroy(em_plan plan) {
  if (!plan) return EM_ERR_INVALID;
  if (plan->stream) (void)hipStreamSynchronize(plan->stream);
  if (plan->gstream) (void)hipStreamSynchronize(plan->gstream);
  delete plan;
  return EM_OK;
}

em_status em_plan_set_stream(em_plan plan, void* s) {
  if (!plan) return EM_ERR_INVALID;
  plan->stream = reinterpret_cast<hipStream_t>(s);
  return EM_OK;
}

em_status em_plan_constants(em_plan plan, double* lambda, double* sigma, double* gamma) {
  if (!plan) return EM_ERR_INVALID;
  if (lambda) *lambda = plan->k.lambda;
  if (sigma) *sigma = plan->k.sigma;
  if (gamma) *gamma = plan->k.gamma;
  return EM_OK;
}

em_status em_plan_set_probe_group(em_plan plan, int g) {
  if (!plan || g < 0) return EM_ERR_INVALID;
  plan->tune.probe_group = g;
  return EM_OK;
}

em_status em_plan_group_size(em_plan plan, int nprobe, int* g) {
  if (!plan || nprobe <= 0 || !g) return EM_ERR_INVALID;
  *g = group_size(plan, nprobe);
  return EM_OK;
}

em_status em_plan_set_pass_timing(em_plan plan, int on) {
  if (!plan) return EM_ERR_INVALID;
  if (plan->gstream) (void)hipStreamSynchronize(plan->gstream);
  plan->clear_pass_events();
  plan->pass_timing = on != 0;
  return EM_OK;
}

em_status em_plan_pass_times(em_plan plan, double* row_ms, double* col_ms, int* n_row, int* n_col) {
  if (!plan || !row_ms || !col_ms || !n_row || !n_col) return EM_ERR_INVALID;
  double sum[2] = {0, 0};
  int cnt[2] = {0, 0};
  em_status st = EM_OK;
  for (size_t i = 0; i < plan->tkind.size() && st == EM_OK; ++i) {
    st = plan->hip(hipEventSynchronize(plan->tev[2 * i + 1]), "event sync");
    float ms = 0;
    if (st == EM_OK) st = plan->hip(hipEventElapsedTime(&ms, plan->tev[2 * i], plan->tev[2 * i + 1]), "elapsed");
    sum[plan->tkind[i]] += ms;
    cnt[plan->tkind[i]] += 1;
  }
  plan->clear_pass_events();
  *row_ms = cnt[0] ? sum[0] / cnt[0] : 0.0;
  *col_ms = cnt[1] ? sum[1] / cnt[1] : 0.0;
  *n_row = cnt[0];
  *n_col = cnt[1];
  return st;
}

em_status em_plan_get_tuning(em_plan plan, em_tuning* t) {
  if (!plan || !t) return EM_ERR_INVALID;
  *t = plan->tune;
  return EM_OK;
}

em_status em_plan_set_tuning(em_plan plan, const em_tuning* t) {
  if (!plan || !t || t->row_blocks <= 0 || t->col_blocks <= 0 || t->probe_group < 0 || t->col_variant < 0 ||
      t->col_variant > 2 || t->row_variant < 0 || t->row_variant > 2 || t->persistent < 0 || t->persistent > 1 ||
      t->row_radix < 0 || t->row_radix >= fft_variants(plan->grid.nx) || t->col_radix < 0 ||
      t->col_radix >= fft_variants(plan->grid.ny))
    return EM_ERR_INVALID;
  if (t->col_tile != 0 && t->col_tile != col_tile(plan->grid.ny) &&
      (t->col_tile != col_alt_tile(plan->grid.ny) || plan->grid.nx % t->col_tile != 0) &&
      !(t->col_tile == 8 && col_transposed_available(plan->grid.ny) && plan->grid.nx % 8 == 0))
    return EM_ERR_INVALID;
  plan->tune = *t;
  return EM_OK;
}

const char* em_plan_last_error(em_plan plan) { return plan ? plan->err.c_str() : "null plan"; }

em_status em_set_element_coefs(em_plan plan, int Z, int nk, const double* a, const double* b, int ng,
                               const double* c, const double* d) {
  if (!plan || Z <= 0 || nk < 0 || ng < 0 || (nk && (!a || !b)) || (ng && (!c || !d)))
    return EM_ERR_INVALID;
  ElementCoefs k;
  k.a.assign(a, a + nk);
  k.b.assign(b, b + nk);
  if (ng) {
    k.c.assign(c, c + ng);
    k.d.assign(d, d + ng);
  }
  plan->custom[Z] = k;
  plan->spline_cache.erase(Z);
  return EM_OK;
}

em_status em_set_spline_params(em_plan plan, int nR, double r_min, double r_max, double tap_frac) {
  if (!plan || nR < 4 || nR > 256 || !(r_min > 0) || !(r_max > r_min) || !(tap_frac > 0) || !(tap_frac < 1))
    return EM_ERR_INVALID;
  plan->nR = nR;
  plan->r_min = r_min;
  plan->r_max = r_max;
  plan->tap = tap_frac;
  plan->spline_cache.clear();
  return EM_OK;
}

em_status em_set_element_params(em_plan plan, int Z, int ptype, const double* cl, const double* cnl) {
  if (!plan || Z <= 0 || Z > 103 || ptype < EM_PT_DOYLE_0_4 || ptype > EM_PT_LOBATO_0_12 || !cl || !cnl)
    return EM_ERR_INVALID;
  em_plan_s::Params3D q;
  q.pt = ptype;
  const int nterms = (ptype == EM_PT_DOYLE_0_4) ? 4 : (ptype == EM_PT_KIRKLAND_0_12 || ptype == EM_PT_WEICKENMEIER_0_12) ? 6 : 5;
  for (int i = 0; i < 6; ++i) {
    q.cl[i] = i < nterms ? cl[i] : 0.0;
    q.cnl[i] = i < nterms ? cnl[i] : 1.0;
    // the forms divide by, or exponentiate with, the non-linear coefficients
    if (i < nterms && !(cnl[i] > 0)) return EM_ERR_INVALID;
  }
  plan->params3d[Z] = q;
  return EM_OK;
}

em_status em_plan_set_fine_slicing(em_plan plan, int enable, int nq) {
  if (!plan || nq < 1 || nq > kMaxFineQ) return EM_ERR_INVALID;
  plan->fine = enable != 0;
  plan->nqz = nq;
  return EM_OK;
}

em_status em_plan_set_specimen(em_plan p, const em_atoms* at, int border, int* n_slices) {
  if (!p || !at || at->n < 0 || (at->n > 0 && (!at->x || !at->y || !at->z || !at->Z))) return EM_ERR_INVALID;
  if (border < 0 || border > 2) return EM_ERR_INVALID;
  const long long n = at->n;
  // sort by z (stable): Slice_Generator requires z-sorted atoms (slice_generator.hpp:55)
  std::vector<long long> order(n);
  for (long long i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](long long a, long long b) { return at->z[a] < at->z[b]; });
  p->ax.resize(n); p->ay.resize(n); p->az.resize(n); p->aocc.resize(n); p->aZ.resize(n);
  for (long long i = 0; i < n; ++i) {
    const long long j = order[i];
    p->ax[i] = at->x[j];
    p->ay[i] = at->y[j];
    p->az[i] = at->z[j];
    p->aZ[i] = at->Z[j];
    p->aocc[i] = at->occ ? at->occ[j] : 1.0;
    if (at->Z[j] <= 0 || at->Z[j] > 103) return p->fail(EM_ERR_INVALID, "atomic number out of range");
  }
  p->types = p->aZ;
  std::sort(p->types.begin(), p->types.end());
  p->types.erase(std::unique(p->types.begin(), p->types.end()), p->types.end());
  if ((int)p->types.size() > 8) return p->fail(EM_ERR_UNSUPPORTED, "more than 8 elements");
  p->splines.clear();
  for (int Z : p->types) {
    auto c = p->spline_cache.find(Z);
    if (c == p->spline_cache.end()) {
      auto it = p->custom.find(Z);
      const ElementCoefs k = it != p->custom.end() ? it->second : moliere_coefs(Z);
      c = p->spline_cache.emplace(Z, build_spline(k, p->nR, p->r_min, p->r_max, p->tap)).first;
    }
    p->splines.push_back(c->second);
  }
  double rmax = 0;
  for (auto& s : p->splines) rmax = std::max(rmax, std::sqrt(s.rmax2));
  p->slices = p->fine ? make_slices_fine(p->az, p->grid.dz, border, kZEps, rmax)
                      : make_slices(p->az, p->grid.dz, border, kZEps);
  const int S = (int)p->slices.size();
  // bins of size >= R_max covering [-bin, lx + bin) x [-bin, ly + bin)
  const double bin = std::max(rmax, 1e-3);
  p->bin = (float)bin;
  p->nbx = (int)std::ceil((p->grid.lx + 2 * bin) / bin);
  p->nby = (int)std::ceil((p->grid.ly + 2 * bin) / bin);
  const int nb = p->nbx * p->nby;
  std::vector<float4> atoms_h;
  std::vector<float> az_h;  // fine slicing: slice centre height above each entry's atom
  std::vector<int> off_h((size_t)std::max(S, 1) * (nb + 1), 0);
  atoms_h.reserve(n);
  for (int s = 0; s < S; ++s) {
    const SliceRange& r = p->slices[s];
    const double zmid = 0.5 * (r.z_start + r.z_end);
    std::vector<std::vector<long long>> bins(nb);
    for (long long i = r.i0; i <= r.i1; ++i) {
      int bx = (int)std::floor((p->ax[i] + bin) / bin), by = (int)std::floor((p->ay[i] + bin) / bin);
      bx = std::min(std::max(bx, 0), p->nbx - 1);
      by = std::min(std::max(by, 0), p->nby - 1);
      bins[by * p->nbx + bx].push_back(i);
    }
    int* o = off_h.data() + (size_t)s * (nb + 1);
    for (int bI = 0; bI < nb; ++bI) {
      o[bI] = (int)atoms_h.size();
      for (long long i : bins[bI]) {
        const int t = (int)(std::lower_bound(p->types.begin(), p->types.end(), p->aZ[i]) - p->types.begin());
        float4 a;
        a.x = (float)p->ax[i];
        a.y = (float)p->ay[i];
        a.z = (float)(p->vr_factor * p->aocc[i]);
        int ti = t;
        std::memcpy(&a.w, &ti, sizeof(int));
        atoms_h.push_back(a);
        if (p->fine) az_h.push_back((float)(zmid - p->az[i]));
      }
    }
    o[nb] = (int)atoms_h.size();
  }
  if (atoms_h.empty()) atoms_h.push_back(make_float4(0, 0, 0, 0));
  const int nt = (int)p->types.size();
  std::vector<float> tab_h((size_t)std::max(nt, 1) * 5 * p->nR, 0.f), rm_h(std::max(nt, 1), 0.f);
  for (int t = 0; t < nt; ++t) {
    const Spline& sp = p->splines[t];
    float* tb = tab_h.data() + (size_t)t * 5 * p->nR;
    for (int i = 0; i < p->nR; ++i) {
      tb[i] = (float)sp.R2[i];
      tb[p->nR + i] = (float)sp.c0[i];
      tb[2 * p->nR + i] = (float)sp.c1[i];
      tb[3 * p->nR + i] = (float)sp.c2[i];
      tb[4 * p->nR + i] = (float)sp.c3[i];
    }
    rm_h[t] = (float)sp.rmax2;
  }
  EM_HIP(p, p->atoms.ensure(sizeof(float4) * atoms_h.size()));
  EM_HIP(p, hipMemcpy(p->atoms.p, atoms_h.data(), sizeof(float4) * atoms_h.size(), hipMemcpyHostToDevice));
  EM_HIP(p, p->off.ensure(sizeof(int) * off_h.size()));
  EM_HIP(p, hipMemcpy(p->off.p, off_h.data(), sizeof(int) * off_h.size(), hipMemcpyHostToDevice));
  EM_HIP(p, p->tables.ensure(sizeof(float) * tab_h.size()));
  EM_HIP(p, hipMemcpy(p->tables.p, tab_h.data(), sizeof(float) * tab_h.size(), hipMemcpyHostToDevice));
  EM_HIP(p, p->rmax2.ensure(sizeof(float) * rm_h.size()));
  EM_HIP(p, hipMemcpy(p->rmax2.p, rm_h.data(), sizeof(float) * rm_h.size(), hipMemcpyHostToDevice));
  if (p->fine && !az_h.empty()) {
    // per-entry splines of the potential integrated over each entry's slice
    const long long ne = (long long)az_h.size();
    std::vector<FineElement> el(nt);
    std::vector<double> r2(std::max<size_t>(1, (size_t)nt * p->nR));
    const double r_tap = p->tap * p->r_max;
    for (int t = 0; t < nt; ++t) {
      auto it = p->params3d.find(p->types[t]);
      FineElement& e = el[t];
      if (it != p->params3d.end()) {
        e.pt = it->second.pt;
        for (int k = 0; k < 6; ++k) {
          e.cl[k] = it->second.cl[k];
          e.cnl[k] = it->second.cnl[k];
        }
      } else {
        e.pt = EM_PT_KIRKLAND_0_12;
        moliere_3d(p->types[t], e.cl, e.cnl);
      }
      e.r2_tap = r_tap * r_tap;
      e.tap_cf = 0.5 * kPi / (p->r_max * p->r_max - e.r2_tap);
      for (int i = 0; i < p->nR; ++i) r2[(size_t)t * p->nR + i] = p->splines[t].R2[i];
    }
    std::vector<double> qx, qw;
    gauss_legendre(p->nqz, qx, qw);
    std::vector<double2> q(p->nqz);
    for (int k = 0; k < p->nqz; ++k) q[k] = make_double2(qx[k], qw[k]);
    EM_HIP(p, p->fine_az.ensure(sizeof(float) * ne));
    EM_HIP(p, hipMemcpy(p->fine_az.p, az_h.data(), sizeof(float) * ne, hipMemcpyHostToDevice));
    EM_HIP(p, p->fine_el.ensure(sizeof(FineElement) * el.size()));
    EM_HIP(p, hipMemcpy(p->fine_el.p, el.data(), sizeof(FineElement) * el.size(), hipMemcpyHostToDevice));
    EM_HIP(p, p->fine_r2.ensure(sizeof(double) * r2.size()));
    EM_HIP(p, hipMemcpy(p->fine_r2.p, r2.data(), sizeof(double) * r2.size(), hipMemcpyHostToDevice));
    EM_HIP(p, p->fine_q.ensure(sizeof(double2) * q.size()));
    EM_HIP(p, hipMemcpy(p->fine_q.p, q.data(), sizeof(double2) * q.size(), hipMemcpyHostToDevice));
    EM_HIP(p, p->fine_tab.ensure(sizeof(float4) * ne * p->nR));
    EM_HIP(p, launch_fine_coefs(ne, p->nR, p->fine_r2.as<double>(), p->fine_el.as<FineElement>(),
                                p->atoms.as<float4>(), p->fine_az.as<float>(), p->grid.dz, p->nqz,
                                p->fine_q.as<double2>(), p->fine_tab.as<float4>(), p->stream));
  }
  p->fine_entries = p->fine ? (long long)az_h.size() : 0;
  p->has_specimen = true;
  p->vstore_slices = -1;
  if (n_slices) *n_slices = S;
  return EM_OK;
}

em_status em_slice_info(em_plan p, int s, double* zs, double* ze, long long* i0, long long* i1) {
  if (!p || !p->has_specimen || s < 0 || s >= (int)p->slices.size()) return EM_ERR_INVALID;
  const SliceRange& r = p->slices[s];
  if (zs) *zs = r.z_start;
  if (ze) *ze = r.z_end;
  if (i0) *i0 = r.i0;
  if (i1) *i1 = r.i1;
  return EM_OK;
}

em_status em_spline_tables(em_plan p, int max_types, int* Zs, int* ntypes, int* nR, double* R2,
                           double* c0, double* c1, double* c2, double* c3) {
  if (!p || !p->has_specimen) return EM_ERR_INVALID;
  const int nt = (int)p->types.size();
  if (ntypes) *ntypes = nt;
  if (nR) *nR = p->nR;
  if (max_types < nt) return EM_OK;  // sizes only
  for (int t = 0; t < nt; ++t) {
    if (Zs) Zs[t] = p->types[t];
    const Spline& s = p->splines[t];
    const size_t o = (size_t)t * p->nR;
    for (int i = 0; i < p->nR; ++i) {
      if (R2) R2[o + i] = s.R2[i];
      if (c0) c0[o + i] = s.c0[i];
      if (c1) c1[o + i] = s.c1[i];
      if (c2) c2[o + i] = s.c2[i];
      if (c3) c3[o + i] = s.c3[i];
    }
  }
  return EM_OK;
}

em_status em_potential_slices(em_plan p, int s0, int ns, float* V) {
  if (!p || !p->has_specimen || !V || s0 < 0 || ns <= 0 || s0 + ns > (int)p->slices.size())
    return EM_ERR_INVALID;
  const int nb = p->nbx * p->nby;
  EM_HIP(p, launch_potential(p->gdev, ns, p->off.as<int>() + (size_t)s0 * (nb + 1), p->nbx, p->nby,
                             p->bin, p->atoms.as<float4>(), (int)p->types.size(), p->nR,
                             p->tables.as<float>(), p->rmax2.as<float>(),
                             p->fine_entries > 0 ? p->fine_tab.as<float4>() : nullptr, V, p->stream));
  return EM_OK;
}

em_status em_propagator(em_plan p, void* P) {
  if (!p || !P) return EM_ERR_INVALID;
  EM_HIP(p, hipMemcpyAsync(P, p->P.p, sizeof(float2) * p->N(), hipMemcpyDeviceToDevice, p->stream));
  return EM_OK;
}

}  // extern "C"

namespace {

// Lens table A for the given lens/defocus, probe ramps for nb probes, then the
// column pass that turns A * ramps into F_y^-1 of the Fourier-space probe
// (the (y, kx) layout the row pass reads).
em_status probes_mixed(em_plan p, const em_lens* cond, double defocus_shift, const double* x,
                       const double* y, int nb, float2* wave) {
  const long long N = p->N();
  EM_HIP(p, p->A.ensure(sizeof(float2) * N));
  EM_HIP(p, launch_lens_table(p->gdev, derive_lens(*cond, p->k.lambda, defocus_shift), p->A.as<float2>(), p->stream));
  std::vector<double> xe(nb), ye(nb);
  for (int b = 0; b < nb; ++b) {
    xe[b] = -2.0 * kPi * (x[b] - 0.5 * p->grid.lx);
    ye[b] = -2.0 * kPi * (y[b] - 0.5 * p->grid.ly);
  }
  EM_HIP(p, p->xe.ensure(sizeof(double) * nb));
  EM_HIP(p, p->ye.ensure(sizeof(double) * nb));
  EM_HIP(p, hipMemcpyAsync(p->xe.p, xe.data(), sizeof(double) * nb, hipMemcpyHostToDevice, p->stream));
  EM_HIP(p, hipMemcpyAsync(p->ye.p, ye.data(), sizeof(double) * nb, hipMemcpyHostToDevice, p->stream));
  EM_HIP(p, p->ex.ensure(sizeof(float2) * nb * p->grid.nx));
  EM_HIP(p, p->ey.ensure(sizeof(float2) * nb * p->grid.ny));
  EM_HIP(p, launch_probe_ramps(p->gdev, nb, p->xe.as<double>(), p->ye.as<double>(), p->ex.as<float2>(),
                               p->ey.as<float2>(), p->stream));
  ColArgs c = col_args(p, nullptr, wave, nb);
  c.A = p->A.as<float2>();
  c.ex = p->ex.as<float2>();
  c.ey = p->ey.as<float2>();
  EM_TRY(col(p, COL_LOAD_PROBE | COL_POST_IFFT | COL_STORE, c));
  // the host vectors must outlive the asynchronous copies
  EM_HIP(p, hipStreamSynchronize(p->stream));
  return EM_OK;
}

}  // namespace

extern "C" {

em_status em_build_probes(em_plan p, const em_lens* cond, double defocus_shift, const double* x,
                          const double* y, int nprobe, void* psi) {
  if (!p || !cond || !x || !y || !psi || nprobe <= 0) return EM_ERR_INVALID;
  float2* w = reinterpret_cast<float2*>(psi);
  EM_TRY(probes_mixed(p, cond, defocus_shift, x, y, nprobe, w));
  RowArgs r = row_args(p, w, w, nprobe);
  EM_TRY(row(p, ROW_PRE_IFFT, r));
  return EM_OK;
}

em_status em_plane_wave(em_plan p, int nprobe, void* psi) {
  if (!p || !psi || nprobe <= 0) return EM_ERR_INVALID;
  EM_HIP(p, launch_fill(reinterpret_cast<float2*>(psi), p->N() * nprobe, make_float2(1.f, 0.f), p->stream));
  return EM_OK;
}

em_status em_fft2d(em_plan p, void* data, int nbatch, int dir) {
  if (!p || !data || nbatch <= 0 || (dir != -1 && dir != 1)) return EM_ERR_INVALID;
  float2* w = reinterpret_cast<float2*>(data);
  if (dir < 0) {
    EM_TRY(row(p, ROW_POST_FFT, row_args(p, w, w, nbatch)));
    EM_TRY(col(p, COL_PRE_FFT | COL_STORE, col_args(p, w, w, nbatch)));
  } else {
    EM_TRY(col(p, COL_PRE_IFFT | COL_STORE, col_args(p, w, w, nbatch)));
    EM_TRY(row(p, ROW_PRE_IFFT, row_args(p, w, w, nbatch)));
  }
  return EM_OK;
}

em_status em_multislice_forward(em_plan p, const float* V, const void* t, int S, void* psi, int nprobe,
                                void* ckpt, int every) {
  if (!p || !psi || S <= 0 || nprobe <= 0 || (!V) == (!t) || every <= 0) return EM_ERR_INVALID;
  float2* w = reinterpret_cast<float2*>(psi);
  const int G = group_size(p, nprobe);
  const bool pk = !ckpt && use_persistent(p, t && !V, S);
  if (pk) EM_TRY(persistent_begin(p));
  for (int g0 = 0; g0 < nprobe; g0 += G) {
    const int nb = std::min(G, nprobe - g0);
    float2* ck = ckpt ? reinterpret_cast<float2*>(ckpt) + (long long)g0 * p->N() : nullptr;
    EM_TRY(forward_group(p, V, reinterpret_cast<const float2*>(t), S, w + (long long)g0 * p->N(), nb, ck,
                         (long long)nprobe * p->N(), every, false));
  }
  if (pk) EM_TRY(persistent_check(p));
  return EM_OK;
}

em_status em_multislice_forward_multi(em_plan p, const float* const* V, const int* S, void* psi, int nprobe) {
  if (!p || !V || !S || !psi || nprobe <= 0) return EM_ERR_INVALID;
  for (int b = 0; b < nprobe; ++b)
    if (!V[b] || S[b] <= 0) return EM_ERR_INVALID;
  const long long N = p->N();
  float2* w = reinterpret_cast<float2*>(psi);
  const float2* P = p->P.as<float2>();
  // Every launch's chunk table is built before any launch is enqueued.  A
  // chunk is a run of consecutive probes that share one potential stack.
  struct Launch {
    bool row;
    unsigned mode;
    int skip_in, skip_out, first, count;
  };
  std::vector<PassChunk> chunks;
  std::vector<Launch> launches;
  auto add = [&](const std::vector<int>& probes, int pb, bool row_table, int s) {
    const int first = (int)chunks.size();
    for (size_t i = 0; i < probes.size();) {
      const int b = probes[i];
      int n = 1;
      while (i + n < probes.size() && n < pb && probes[i + n] == b + n && V[b + n] == V[b]) ++n;
      PassChunk ch{};
      ch.out = w + (long long)b * N;
      ch.in = ch.out;
      ch.table = row_table ? static_cast<const void*>(V[b] + (long long)s * N) : static_cast<const void*>(P);
      ch.n = n;
      chunks.push_back(ch);
      i += n;
    }
    return std::make_pair(first, (int)chunks.size() - first);
  };
  const int blocks = p->grid.nx / col_tile(p->grid.ny);
  auto col_chunk_pb = [&](int btot) {
    const int groups = std::max(1, std::min(btot, (std::max(1, p->tune.col_blocks) + blocks - 1) / blocks));
    return (btot + groups - 1) / groups;
  };
  const int G = group_size(p, nprobe);
  for (int g0 = 0; g0 < nprobe; g0 += G) {
    const int g1 = std::min(nprobe, g0 + G);
    int Smax = 0;
    for (int b = g0; b < g1; ++b) Smax = std::max(Smax, S[b]);
    for (int s = 0; s < Smax; ++s) {
      // row pass: output columns below the band-skip threshold are dropped
      // except on a probe's last slice (its final column transform keeps all)
      for (int so = 0; so < 2; ++so) {
        std::vector<int> probes;
        for (int b = g0; b < g1; ++b)
          if (s < S[b] && (s < S[b] - 1 ? 1 : 0) == so) probes.push_back(b);
        if (probes.empty()) continue;
        const auto fc = add(probes, row_pb(p, (int)probes.size()), true, s);
        launches.push_back({true, (s > 0 ? ROW_PRE_IFFT : 0u) | ROW_MUL_V | ROW_POST_FFT | ROW_CHUNKS, s > 0, so,
                            fc.first, fc.second});
      }
      // column pass: the propagator step, or the last slice's FFT_y
      for (int last = 0; last < 2; ++last) {
        std::vector<int> probes;
        for (int b = g0; b < g1; ++b)
          if (last ? s == S[b] - 1 : s < S[b] - 1) probes.push_back(b);
        if (probes.empty()) continue;
        const auto fc = add(probes, col_chunk_pb((int)probes.size()), false, s);
        launches.push_back({false,
                            last ? (COL_PRE_FFT | COL_STORE | COL_CHUNKS)
                                 : (COL_PRE_FFT | COL_MUL_P | COL_POST_IFFT | COL_STORE | COL_CHUNKS),
                            0, 0, fc.first, fc.second});
      }
    }
  }
  for (const Launch& l : launches)
    if (l.count > 65535) return p->fail(EM_ERR_INVALID, "too many probes in one pass");
  // an earlier call's launches may still be reading the table
  EM_HIP(p, hipStreamSynchronize(p->stream));
  EM_HIP(p, p->mchunks.ensure(sizeof(PassChunk) * chunks.size()));
  EM_HIP(p, hipMemcpy(p->mchunks.p, chunks.data(), sizeof(PassChunk) * chunks.size(), hipMemcpyHostToDevice));
  const PassChunk* dchunks = p->mchunks.as<PassChunk>();
  for (const Launch& l : launches) {
    if (l.row) {
      RowArgs ra = row_args(p, nullptr, nullptr, l.count);
      ra.pb = 1;
      ra.chunks = dchunks + l.first;
      ra.skip_in = l.skip_in;
      ra.skip_out = l.skip_out;
      ra.zk = (l.skip_in && l.skip_out) ? p->zk_x : 0;
      EM_TRY(row(p, l.mode, ra));
    } else {
      ColArgs cc = col_args(p, nullptr, nullptr, l.count);
      cc.pb = 1;
      cc.chunks = dchunks + l.first;
      cc.cols = 0;
      cc.variant = COL_VARIANT_BLOCK;
      cc.skip_lo = p->skip_lo;
      cc.skip_hi = p->skip_hi;
      cc.zk = p->zk_y;
      EM_TRY(col(p, l.mode, cc));
    }
  }
  return EM_OK;
}

}  // extern "C"

namespace {

// Backward pass of one probe group (nb probes at lam; their checkpoints at ck,
// slice j of them at ck + j * ck_stride).  The gradient of each slice sums
// this group's per-workgroup partials; `first` overwrites gt, later groups
// add to it, so the result does not depend on scheduling.
em_status adjoint_group(em_plan p, const float* V, const float2* tc, int S, const float2* ck, long long ck_stride,
                        int every, float2* lam, int nb, float2* gtc, float2* gpsi0, bool first) {
  const long long N = p->N();
  const long long GN = (long long)nb * N;
  const int pb = row_pb(p, nb);
  const int nparts = (nb + pb - 1) / pb;
  const bool direct = first && nparts == 1;  // the single partial is the gradient
  if (!direct) EM_HIP(p, p->gpart.ensure(sizeof(float2) * N * nparts));
  if (every > 1) {
    // segment workspace: the recomputed waves entering each slice of a segment
    EM_HIP(p, p->seg.ensure(sizeof(float2) * GN * every));
    EM_HIP(p, p->segwave.ensure(sizeof(float2) * GN));
  }
  // lambda_{S-1} = FFT^H gPsi: column inverse now, row inverse inside the first backward row pass
  EM_TRY(col(p, COL_PRE_IFFT | COL_STORE, col_args(p, lam, lam, nb)));
  for (int k = S - 1; k >= 0; --k) {
    const float2* ckk;
    if (every == 1) {
      ckk = ck + (long long)k * ck_stride;
    } else {
      const int j = k / every, s0 = j * every;
      if (k == S - 1 || k % every == every - 1) {
        // recompute the waves entering slices s0..k from checkpoint j
        float2* sw = p->segwave.as<float2>();
        float2* sg = p->seg.as<float2>();
        EM_HIP(p, hipMemcpyAsync(sw, ck + (long long)j * ck_stride, sizeof(float2) * GN, hipMemcpyDeviceToDevice,
                                 p->stream));
        for (int s = s0; s <= k; ++s) {
          RowArgs r = row_args_skip(p, sw, sw, nb, s > s0, s < k);
          r.ckpt = sg + (long long)(s - s0) * GN;
          unsigned m = ROW_CKPT_STORE | (s > s0 ? ROW_PRE_IFFT : 0u);
          if (s < k) {
            m |= ROW_POST_FFT | (V ? ROW_MUL_V : ROW_MUL_T);
            if (V) r.V = V + (long long)s * N; else r.t = tc + (long long)s * N;
          }
          EM_TRY(row(p, m, r));
          if (s < k) EM_TRY(col(p, COL_PRE_FFT | COL_MUL_P | COL_POST_IFFT | COL_STORE, col_args_skip(p, sw, sw, nb)));
        }
      }
      ckk = p->seg.as<float2>() + (long long)(k - s0) * GN;
    }
    RowArgs r = row_args_skip(p, lam, lam, nb, k < S - 1, k > 0);
    r.ckpt = const_cast<float2*>(ckk);
    r.gpart = direct ? gtc + (long long)k * N : p->gpart.as<float2>();
    unsigned m = ROW_PRE_IFFT | ROW_GRAD;
    if (k > 0 || gpsi0) {
      m |= ROW_CONJ | (V ? ROW_MUL_V : ROW_MUL_T);
      if (V) r.V = V + (long long)k * N; else r.t = tc + (long long)k * N;
    }
    if (k > 0) m |= ROW_POST_FFT;
    else if (gpsi0) r.out = gpsi0;  // conj(t_0) lambda_0, real space
    EM_TRY(row(p, m, r));
    if (!direct)
      EM_HIP(p, launch_grad_reduce(N, nparts, p->gpart.as<float2>(), gtc + (long long)k * N, first ? 0 : 1,
                                   p->stream));
    if (k > 0) EM_TRY(col(p, COL_PRE_FFT | COL_CONJ_P | COL_POST_IFFT | COL_STORE, col_args_skip(p, lam, lam, nb)));
  }
  return EM_OK;
}

}  // namespace

extern "C" {

em_status em_multislice_adjoint(em_plan p, const float* V, const void* t, int S, const void* ckpt, int every,
                                void* gPsi, int nprobe, void* gt, float* gV, void* gpsi0) {
  if (!p || !ckpt || !gPsi || !gt || S <= 0 || nprobe <= 0 || (!V) == (!t) || every <= 0) return EM_ERR_INVALID;
  const long long N = p->N();
  const long long BN = (long long)nprobe * N;
  float2* lam = reinterpret_cast<float2*>(gPsi);
  const float2* ck = reinterpret_cast<const float2*>(ckpt);
  const float2* tc = reinterpret_cast<const float2*>(t);
  float2* gtc = reinterpret_cast<float2*>(gt);
  float2* g0 = reinterpret_cast<float2*>(gpsi0);
  // probe groups as in the forward pass: a group's lambdas and checkpoints
  // stay near the last-level cache for the whole backward sweep
  const int G = group_size(p, nprobe);
  for (int b0 = 0; b0 < nprobe; b0 += G) {
    const int nb = std::min(G, nprobe - b0);
    EM_TRY(adjoint_group(p, V, tc, S, ck + (long long)b0 * N, BN, every, lam + (long long)b0 * N, nb, gtc,
                         g0 ? g0 + (long long)b0 * N : nullptr, b0 == 0));
  }
  if (V && gV) EM_HIP(p, launch_grad_v(gtc, V, 1.f, N * S, gV, p->stream));
  return EM_OK;
}

em_status em_detector_sums(em_plan p, const void* Psi, int nprobe, int ndet, const double* inner,
                           const double* outer, float w, int accumulate, float* out) {
  if (!p || !Psi || !out || nprobe <= 0 || ndet <= 0 || ndet > kMaxDet || !inner || !outer)
    return EM_ERR_INVALID;
  const int ntiles = col_ntiles(p->grid.nx, p->grid.ny);
  EM_HIP(p, p->dpart.ensure(sizeof(float) * (size_t)nprobe * ndet * ntiles));
  ColArgs c = col_args(p, reinterpret_cast<const float2*>(Psi), nullptr, nprobe);
  c.ndet = ndet;
  for (int d = 0; d < ndet; ++d) {
    c.dg2min[d] = mrad_to_g2(inner[d], p->k.lambda, false);
    c.dg2max[d] = mrad_to_g2(outer[d], p->k.lambda, true);
  }
  c.dpart = p->dpart.as<float>();
  EM_TRY(col(p, COL_DET, c));
  EM_HIP(p, launch_det_reduce(nprobe, ndet, ntiles, p->dpart.as<float>(), w, accumulate, out, nprobe, p->stream));
  return EM_OK;
}

// em_detector_sums with its per-tile partials in the device's fixed scratch:
// no allocation.  The detector pass writes every partial it reduces, so the
// scratch needs no clearing; calls that overlap on one device share it.
em_status em_detector_sums_noalloc(em_plan p, const void* Psi, int nprobe, int ndet, const double* inner,
                                   const double* outer, float w, int accumulate, float* out) {
  if (!p || !Psi || !out || nprobe <= 0 || ndet <= 0 || ndet > kMaxDet || !inner || !outer)
    return EM_ERR_INVALID;
  if (!p->scratch) return p->fail(EM_ERR_HIP, "no device scratch");
  int dev = -1;
  EM_HIP(p, hipGetDevice(&dev));
  if (dev != p->device) return p->fail(EM_ERR_INVALID, "the plan's device is not current");
  const int ntiles = col_ntiles(p->grid.nx, p->grid.ny);
  if ((size_t)nprobe * ndet * ntiles > kNoAllocScratchFloats)
    return p->fail(EM_ERR_INVALID, "nprobe * ndet * (nx / 8) exceeds the fixed scratch of " +
                                       std::to_string(kNoAllocScratchFloats) + " partial sums");
  ColArgs c = col_args(p, reinterpret_cast<const float2*>(Psi), nullptr, nprobe);
  c.ndet = ndet;
  for (int d = 0; d < ndet; ++d) {
    c.dg2min[d] = mrad_to_g2(inner[d], p->k.lambda, false);
    c.dg2max[d] = mrad_to_g2(outer[d], p->k.lambda, true);
  }
  c.dpart = p->scratch;
  EM_TRY(col(p, COL_DET, c));
  EM_HIP(p, launch_det_reduce(nprobe, ndet, ntiles, p->scratch, w, accumulate, out, nprobe, p->stream));
  return EM_OK;
}

}  // extern "C"

namespace {

// Image of Psi through the lens table in p->A (apply_CTF_kernel +
// inverse FFT + wave_amplitude / wave_spliter): column pass Psi * A then
// IFFT_y, row pass IFFT_x then w |.|^2 or w (re, im), stored or accumulated.
em_status image_from_table(em_plan p, const void* Psi, int nprobe, float* I, float* re, float* im, float w,
                           int accumulate) {
  EM_HIP(p, p->tmp.ensure(sizeof(float2) * p->N() * nprobe));
  ColArgs c = col_args(p, reinterpret_cast<const float2*>(Psi), p->tmp.as<float2>(), nprobe);
  c.A = p->A.as<float2>();
  EM_TRY(col(p, COL_LOAD_MUL_A | COL_POST_IFFT | COL_STORE, c));
  RowArgs r = row_args(p, p->tmp.as<float2>(), nullptr, nprobe);
  r.wgt = w;
  unsigned m = ROW_PRE_IFFT | (accumulate ? ROW_OUT_ACCUM : 0u);
  if (I) {
    r.o1 = I;
    m |= ROW_OUT_ABS2;
  } else {
    r.o1 = re;
    r.o2 = im;
    m |= ROW_OUT_SPLIT;
  }
  return row(p, m, r);
}

bool coherence_valid(const em_coherence* c, bool tcc) {
  if (!c || !(c->dsf_sigma >= 0) || !(c->ssf_beta >= 0) || c->mode < EM_COH_TEMPORAL_SPATIAL ||
      c->mode > EM_COH_SPATIAL)
    return false;
  if (tcc) {
    const bool t = c->mode != EM_COH_SPATIAL, s = c->mode != EM_COH_TEMPORAL;
    if (t && (c->dsf_npoints < 1 || c->dsf_npoints > 64)) return false;
    if (s && (c->ssf_npoints < 1 || c->ssf_npoints > 64)) return false;
  }
  return true;
}

}  // namespace

extern "C" {

em_status em_ctf_image(em_plan p, const em_lens* obj, const void* Psi, int nprobe, float* I, float* re,
                       float* im, float w, int accumulate) {
  if (!p || !obj || !Psi || nprobe <= 0 || (!I && !(re && im))) return EM_ERR_INVALID;
  EM_HIP(p, p->A.ensure(sizeof(float2) * p->N()));
  EM_HIP(p, launch_lens_table(p->gdev, derive_lens(*obj, p->k.lambda, 0.0), p->A.as<float2>(), p->stream));
  return image_from_table(p, Psi, nprobe, I, re, im, w, accumulate);
}

em_status em_pctf_image(em_plan p, const em_lens* obj, const em_coherence* coh, const void* Psi, int nprobe,
                        float* I, float* re, float* im, float w, int accumulate) {
  if (!p || !obj || !Psi || nprobe <= 0 || (!I && !(re && im)) || !coherence_valid(coh, false))
    return EM_ERR_INVALID;
  LensDev l = derive_lens(*obj, p->k.lambda, 0.0);
  // PCTF_TEM drops the spread that the mode leaves out (lens_modulation.cuh:95-102)
  l.envelope = 1;
  l.env_lambda = p->k.lambda;
  l.env_c10 = obj->c_10;
  l.env_c30 = obj->c_30;
  l.env_sigma = coh->mode == EM_COH_SPATIAL ? 0.0 : coh->dsf_sigma;
  l.env_beta = coh->mode == EM_COH_TEMPORAL ? 0.0 : coh->ssf_beta * 1e-3;
  EM_HIP(p, p->A.ensure(sizeof(float2) * p->N()));
  EM_HIP(p, launch_lens_table(p->gdev, l, p->A.as<float2>(), p->stream));
  return image_from_table(p, Psi, nprobe, I, re, im, w, accumulate);
}

em_status em_tcc_image(em_plan p, const em_lens* obj, const em_coherence* coh, const void* Psi, int nprobe,
                       float* I, float* re, float* im, float w, int accumulate) {
  if (!p || !obj || !Psi || nprobe <= 0 || (!I && !(re && im)) || !coherence_valid(coh, true))
    return EM_ERR_INVALID;
  const bool temporal = coh->mode != EM_COH_SPATIAL, spatial = coh->mode != EM_COH_TEMPORAL;
  std::vector<double> tx{0.0}, tw{1.0}, sx{0.0}, sw{1.0};
  if (temporal) {
    gauss_hermite(coh->dsf_npoints, tx, tw);
    for (double& v : tw) v /= std::sqrt(kPi);
  }
  if (spatial) {
    gauss_hermite(coh->ssf_npoints, sx, sw);
    for (double& v : sw) v /= std::sqrt(kPi);
  }
  const double q0 = std::sin(coh->ssf_beta * 1e-3) / p->k.lambda;
  EM_HIP(p, p->A.ensure(sizeof(float2) * p->N()));
  bool first = true;
  for (size_t a = 0; a < sx.size(); ++a)
    for (size_t b = 0; b < sx.size(); ++b)
      for (size_t j = 0; j < tx.size(); ++j) {
        // TCC_TEM's node loops: defocus set per temporal node, tilt per source node
        em_lens lj = *obj;
        lj.c_10 = obj->c_10 + std::sqrt(2.0) * coh->dsf_sigma * tx[j];
        LensDev l = derive_lens(lj, p->k.lambda, 0.0);
        l.gxs = q0 * sx[a];
        l.gys = q0 * sx[b];
        EM_HIP(p, launch_lens_table(p->gdev, l, p->A.as<float2>(), p->stream));
        const float wn = (float)(w * sw[a] * sw[b] * tw[j]);
        EM_TRY(image_from_table(p, Psi, nprobe, I, re, im, wn, first ? accumulate : 1));
        first = false;
      }
  return EM_OK;
}

em_status em_diffraction(em_plan p, const void* Psi, int nprobe, float* I, float w, int accumulate) {
  if (!p || !Psi || !I || nprobe <= 0) return EM_ERR_INVALID;
  const long long N = p->N();
  EM_HIP(p, p->tmp.ensure(sizeof(float) * N * nprobe));
  EM_HIP(p, launch_abs2(reinterpret_cast<const float2*>(Psi), p->tmp.as<float>(), N * nprobe, 1.f, 0, p->stream));
  EM_HIP(p, launch_fft_shift(p->tmp.as<float>(), I, p->grid.nx, p->grid.ny, nprobe, w, accumulate, p->stream));
  return EM_OK;
}

}  // extern "C"

namespace {

// Sigma-scaled potentials of every slice of the current specimen, built once
// and kept (the stored-slice cache of transmission_function.cuh:65-101).
em_status ensure_specimen_potential(em_plan p) {
  if (!p->has_specimen) return p->fail(EM_ERR_INVALID, "no specimen");
  const int S = (int)p->slices.size();
  if (S <= 0) return p->fail(EM_ERR_INVALID, "specimen has no slices");
  if (p->vstore_slices == S && p->Vstore.p) return EM_OK;
  EM_HIP(p, p->Vstore.ensure(sizeof(float) * p->N() * S));
  p->vstore_slices = -1;
  EM_TRY(em_potential_slices(p, 0, S, p->Vstore.as<float>()));
  // transmissions of the stored slices: one sincos per pixel and slice, not
  // one per pixel, slice and workgroup of the row passes
  EM_HIP(p, p->Tstore.ensure(sizeof(float2) * p->N() * S));
  EM_HIP(p, launch_transmission(p->Vstore.as<float>(), p->Tstore.as<float2>(), p->N() * S, p->stream));
  p->vstore_slices = S;
  return EM_OK;
}

// The body of one STEM probe batch (scan_simulator.cuh:88-109): probes of the
// batch (their position ramps already in ex / ey) are processed in groups;
// each group builds its probes, runs every slice and leaves detector partials
// in dpart.
em_status stem_batch_body(em_plan p, int nb, int S, const double* dg2min, const double* dg2max, int ndet) {
  const long long N = p->N();
  const int ntiles = col_ntiles(p->grid.nx, p->grid.ny);
  const int G = group_size(p, nb);
  float2* wave = p->wave.as<float2>();
  const bool pk = use_persistent(p, p->tune.row_variant != 2, S);
  for (int g0 = 0; g0 < nb; g0 += G) {
    const int ng = std::min(G, nb - g0);
    ColArgs c = col_args(p, nullptr, wave, ng);
    c.A = p->A.as<float2>();
    c.ex = p->ex.as<float2>() + (long long)g0 * p->grid.nx;
    c.ey = p->ey.as<float2>() + (long long)g0 * p->grid.ny;
    EM_TRY(col(p, COL_LOAD_PROBE | COL_POST_IFFT | COL_STORE, c));
    for (int s = 0; s < S; ++s) {
      if (pk && s > 0 && s < S - 1) continue;  // in the persistent launch
      // slice 0 reads every column of the freshly built probe (the condenser
      // aperture may reach past the skipped band); later slices read the
      // output of a propagator pass, zero there
      RowArgs r = row_args_skip(p, wave, wave, ng, s > 0, s < S - 1);
      unsigned rm;
      if (p->tune.row_variant == 2) {
        // the potential slice (4 B per pixel) with sincos once per workgroup
        r.V = p->Vstore.as<float>() + (long long)s * N;
        rm = ROW_PRE_IFFT | ROW_MUL_V | ROW_POST_FFT;
      } else {
        r.t = p->Tstore.as<float2>() + (long long)s * N;
        rm = row_t_mode(p, ROW_PRE_IFFT | ROW_MUL_T | ROW_POST_FFT);
      }
      EM_TRY(timed(p, 0, [&] { return row(p, rm, r); }));
      if (pk && s == 0) {
        EM_TRY(persistent_interior(p, p->Tstore.as<float2>(), S, wave, ng));
        continue;
      }
      ColArgs cc = col_args_skip(p, wave, wave, ng);
      if (s < S - 1) {
        EM_TRY(timed(p, 1, [&] { return col(p, COL_PRE_FFT | COL_MUL_P | COL_POST_IFFT | COL_STORE, cc); }));
      } else {
        cc.ndet = ndet;
        for (int d = 0; d < ndet; ++d) {
          cc.dg2min[d] = dg2min[d];
          cc.dg2max[d] = dg2max[d];
        }
        cc.dpart = p->dpart.as<float>() + (long long)g0 * ndet * ntiles;
        EM_TRY(col(p, COL_PRE_FFT | COL_DET, cc));
      }
    }
  }
  return EM_OK;
}

struct StreamSwap {
  em_plan p;
  hipStream_t saved;
  StreamSwap(em_plan p_, hipStream_t s) : p(p_), saved(p_->stream) { p->stream = s; }
  ~StreamSwap() { p->stream = saved; }
};

}  // namespace

extern "C" {

em_status em_stem(em_plan p, const em_lens* cond, int npos, const double* x, const double* y, int ndet,
                  const double* inner, const double* outer, float w, int accumulate, float* out) {
  if (!p || !cond || !x || !y || !out || npos <= 0 || ndet <= 0 || ndet > kMaxDet || !inner || !outer)
    return EM_ERR_INVALID;
  EM_TRY(ensure_specimen_potential(p));
  const int S = (int)p->slices.size();
  const long long N = p->N();
  const int B = std::min(p->max_batch, npos);
  const int ntiles = col_ntiles(p->grid.nx, p->grid.ny);
  double dg2min[kMaxDet], dg2max[kMaxDet];
  for (int d = 0; d < ndet; ++d) {
    dg2min[d] = mrad_to_g2(inner[d], p->k.lambda, false);
    dg2max[d] = mrad_to_g2(outer[d], p->k.lambda, true);
  }
  // every buffer the graphs touch is sized before capture
  EM_HIP(p, p->wave.ensure(sizeof(float2) * N * B));
  EM_HIP(p, p->dpart.ensure(sizeof(float) * (size_t)B * ndet * ntiles));
  EM_HIP(p, p->A.ensure(sizeof(float2) * N));
  EM_HIP(p, p->ex.ensure(sizeof(float2) * B * p->grid.nx));
  EM_HIP(p, p->ey.ensure(sizeof(float2) * B * p->grid.ny));
  EM_HIP(p, p->pos_dev.ensure(sizeof(double) * 2 * (size_t)npos));
  // Probe phase ramps of every position (xe = -2 pi (x_b - x_l/2),
  // grid_dev.cuh:104-111) staged in pinned memory and uploaded once per
  // call; batches read them at their offset.  The previous call's upload must
  // have left the staging buffer first.
  EM_HIP(p, hipEventSynchronize(p->pos_copied));
  EM_HIP(p, p->pos_host.ensure(sizeof(double) * 2 * (size_t)npos));
  double* ph = reinterpret_cast<double*>(p->pos_host.p);
  for (int b = 0; b < npos; ++b) {
    ph[b] = -2.0 * kPi * (x[b] - 0.5 * p->grid.lx);
    ph[npos + b] = -2.0 * kPi * (y[b] - 0.5 * p->grid.ly);
  }
  const double* xe = p->pos_dev.as<double>();
  const double* ye = xe + npos;
  // the private graph stream runs after the caller's stream, and the
  // caller's stream after it: no host synchronisation in the call
  const hipStream_t user = p->stream;
  // the persistent slice loop runs eagerly (a cooperative launch), and the
  // call waits for it to check its barriers
  const bool pk = use_persistent(p, p->tune.row_variant != 2, S);
  EM_HIP(p, hipEventRecord(p->stem_in, user));
  EM_HIP(p, hipStreamWaitEvent(p->gstream, p->stem_in, 0));
  em_status st = EM_OK;
  {
    StreamSwap sw(p, p->gstream);
    st = p->hip(hipMemcpyAsync(p->pos_dev.p, p->pos_host.p, sizeof(double) * 2 * (size_t)npos,
                               hipMemcpyHostToDevice, p->stream), "positions");
    if (st == EM_OK) st = p->hip(hipEventRecord(p->pos_copied, p->stream), "record");
    // condenser-lens table with the defocus moved to the middle of slice 0
    // (incident_wave.cuh:33-35)
    const double zmid0 = 0.5 * (p->slices[0].z_start + p->slices[0].z_end);
    if (st == EM_OK)
      st = p->hip(launch_lens_table(p->gdev, derive_lens(*cond, p->k.lambda, zmid0), p->A.as<float2>(), p->stream),
                  "lens table");
    if (st == EM_OK && pk) st = persistent_begin(p);
    for (int start = 0; start < npos && st == EM_OK; start += B) {
      const int nb = std::min(B, npos - start);
      if ((st = p->hip(launch_probe_ramps(p->gdev, nb, xe + start, ye + start, p->ex.as<float2>(),
                                          p->ey.as<float2>(), p->stream), "probe ramps")) != EM_OK)
        break;
      if (p->pass_timing || pk) {
        // timed launches run eagerly: events between the passes need a plain stream
        if ((st = stem_batch_body(p, nb, S, dg2min, dg2max, ndet)) != EM_OK) break;
      } else {
        StemGraphKey key;
        std::memset(&key, 0, sizeof(key));
        key.nb = nb;
        key.S = S;
        key.ndet = ndet;
        key.skip_lo = p->skip_lo;
        key.skip_hi = p->skip_hi;
        key.skip_lo_y = p->skip_lo_y;
        key.skip_hi_y = p->skip_hi_y;
        key.tune = p->tune;
        for (int d = 0; d < ndet; ++d) {
          key.dg2min[d] = dg2min[d];
          key.dg2max[d] = dg2max[d];
        }
        const void* bufs[10] = {p->wave.p, p->dpart.p, p->A.p, p->ex.p, p->ey.p, p->Tstore.p,
                                p->Vstore.p, p->P.p, p->tw_x.p, p->tw_y.p};
        std::memcpy(key.buf, bufs, sizeof(bufs));
        Graph* graph = nullptr;
        for (auto& e : p->stem_graphs)
          if (e.first == key) graph = e.second.get();
        if (!graph) {
          if (p->stem_graphs.size() >= 8) p->stem_graphs.erase(p->stem_graphs.begin());
          auto g = std::make_unique<Graph>();
          if ((st = p->hip(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal), "begin capture")) != EM_OK)
            break;
          const em_status body = stem_batch_body(p, nb, S, dg2min, dg2max, ndet);
          const hipError_t ec = hipStreamEndCapture(p->stream, &g->g);
          if (body != EM_OK) {
            st = body;
            break;
          }
          if ((st = p->hip(ec, "end capture")) != EM_OK) break;
          if ((st = p->hip(hipGraphInstantiate(&g->x, g->g, nullptr, nullptr, 0), "instantiate")) != EM_OK) break;
          graph = g.get();
          p->stem_graphs.emplace_back(key, std::move(g));
        }
        if ((st = p->hip(hipGraphLaunch(graph->x, p->stream), "graph launch")) != EM_OK) break;
      }
      st = p->hip(launch_det_reduce(nb, ndet, ntiles, p->dpart.as<float>(), w, accumulate, out + start, npos, p->stream),
                  "detector reduce");
    }
    if (st == EM_OK && pk) st = persistent_check(p);
  }
  (void)hipEventRecord(p->stem_out, p->gstream);
  (void)hipStreamWaitEvent(user, p->stem_out, 0);
  return st;
}

em_status em_thickness(em_plan p, const float* V, int S, void* psi, int nprobe, double inner, double outer,
                       int stride, float w, int accumulate, float* frames, int* n_frames) {
  if (!p || !V || !psi || S <= 0 || nprobe <= 0 || !frames) return EM_ERR_INVALID;
  const long long N = p->N();
  float2* wv = reinterpret_cast<float2*>(psi);
  return thickness_loop(p, S, wv, nprobe, inner, outer, stride, w, accumulate, frames, n_frames, [&](int s) {
    RowArgs r = row_args(p, wv, wv, nprobe);
    r.V = V + (long long)s * N;
    return row(p, (s > 0 ? ROW_PRE_IFFT : 0u) | ROW_MUL_V | ROW_POST_FFT, r);
  });
}

em_status em_thickness_multi(em_plan p, const float* const* V, int S, void* psi, int nprobe, double inner,
                             double outer, int stride, float w, int accumulate, float* frames, int* n_frames) {
  if (!p || !V || !psi || S <= 0 || nprobe <= 0 || !frames) return EM_ERR_INVALID;
  for (int b = 0; b < nprobe; ++b)
    if (!V[b]) return EM_ERR_INVALID;
  const long long N = p->N();
  float2* wv = reinterpret_cast<float2*>(psi);
  // chunk table of every slice's row pass: runs of consecutive probes that
  // share one potential stack, each with that stack's slice
  const int pb = row_pb(p, nprobe);
  std::vector<PassChunk> chunks;
  for (int s = 0; s < S; ++s)
    for (int b = 0; b < nprobe;) {
      int n = 1;
      while (b + n < nprobe && n < pb && V[b + n] == V[b]) ++n;
      PassChunk ch{};
      ch.out = wv + (long long)b * N;
      ch.in = ch.out;
      ch.table = V[b] + (long long)s * N;
      ch.n = n;
      chunks.push_back(ch);
      b += n;
    }
  const int per_slice = (int)chunks.size() / S;
  if (per_slice > 65535) return p->fail(EM_ERR_INVALID, "too many probes in one pass");
  // an earlier call's launches may still be reading the table
  EM_HIP(p, hipStreamSynchronize(p->stream));
  EM_HIP(p, p->mchunks.ensure(sizeof(PassChunk) * chunks.size()));
  EM_HIP(p, hipMemcpy(p->mchunks.p, chunks.data(), sizeof(PassChunk) * chunks.size(), hipMemcpyHostToDevice));
  const PassChunk* dchunks = p->mchunks.as<PassChunk>();
  return thickness_loop(p, S, wv, nprobe, inner, outer, stride, w, accumulate, frames, n_frames, [&](int s) {
    RowArgs r = row_args(p, nullptr, nullptr, per_slice);
    r.pb = 1;
    r.chunks = dchunks + (long long)s * per_slice;
    return row(p, (s > 0 ? ROW_PRE_IFFT : 0u) | ROW_MUL_V | ROW_POST_FFT | ROW_CHUNKS, r);
  });
}

em_status em_multislice_ragged(int nreq, em_plan* plans, const float* const* V, const int* S, void* const* psi,
                               const int* nprobe, int ndet, const double* inner, const double* outer, int stride,
                               float* const* out, void* hip_stream) {
  if (nreq <= 0 || !plans || !V || !S || !psi || !nprobe || !out || ndet <= 0 || ndet > kMaxDet || !inner ||
      !outer || stride <= 0)
    return EM_ERR_INVALID;
  for (int r = 0; r < nreq; ++r) {
    if (!plans[r] || !V[r] || !psi[r] || S[r] <= 0 || nprobe[r] <= 0 || !out[r]) return EM_ERR_INVALID;
    for (int q = 0; q < r; ++q)
      if (plans[q] == plans[r]) return plans[0]->fail(EM_ERR_INVALID, "requests must use distinct plans");
  }
  hipStream_t user = reinterpret_cast<hipStream_t>(hip_stream);
  em_plan p0 = plans[0];

  // Classes: requests of one grid size and band-skip range step through their
  // slices together, and every fused pass of a slice step is one launch over
  // all of their probes (a chunk table gives each workgroup row its request's
  // wave and V or P).  Detector passes stay per request (their bands depend
  // on the request's wavelength).
  std::vector<std::vector<int>> cls;
  for (int r = 0; r < nreq; ++r) {
    em_plan p = plans[r];
    size_t c = 0;
    for (; c < cls.size(); ++c) {
      em_plan q = plans[cls[c][0]];
      if (q->grid.nx == p->grid.nx && q->grid.ny == p->grid.ny && q->skip_lo == p->skip_lo &&
          q->skip_hi == p->skip_hi && q->skip_lo_y == p->skip_lo_y && q->skip_hi_y == p->skip_hi_y)
        break;
    }
    if (c == cls.size()) cls.emplace_back();
    cls[c].push_back(r);
  }
  // The launches of every class in order, with their chunk tables.
  enum OpKind { OP_ROW, OP_DET, OP_COL };
  struct Op {
    OpKind kind;
    int first, count;  // chunk range (OP_ROW, OP_COL)
    unsigned mode;     // OP_ROW
    int skip_in, skip_out;
    int req, io;       // OP_DET
  };
  std::vector<PassChunk> chunks;
  std::vector<std::vector<Op>> ops(cls.size());
  for (size_t c = 0; c < cls.size(); ++c) {
    em_plan pc = plans[cls[c][0]];
    const long long N = pc->N();
    int Smax = 0;
    for (int r : cls[c]) Smax = std::max(Smax, S[r]);
    std::vector<int> io(nreq, 0);
    auto add_chunks = [&](const std::vector<int>& reqs, int pb, bool row_table, int s) {
      const int first = (int)chunks.size();
      for (int r : reqs)
        for (int j = 0; j < nprobe[r]; j += pb) {
          PassChunk ch{};
          ch.out = reinterpret_cast<float2*>(psi[r]) + (long long)j * N;
          ch.in = ch.out;
          ch.table = row_table ? static_cast<const void*>(V[r] + (long long)s * N)
                               : static_cast<const void*>(plans[r]->P.as<float2>());
          ch.n = std::min(pb, nprobe[r] - j);
          chunks.push_back(ch);
        }
      return std::make_pair(first, (int)chunks.size() - first);
    };
    // a class whose one chunk of row workgroups fills the resident slots
    // (2048 and up) takes all its probes per workgroup: each row's
    // transmission is then formed once per slice step (ragged mix: a 2048^2
    // request of 2 probes, +4%)
    const bool one_row_chunk = pc->grid.ny / row_lines_per_block(pc->grid.nx) >= row_slots(pc);
    for (int s = 0; s < Smax; ++s) {
      const bool rec_step = (s + 1) % stride == 0;
      // row pass: one launch per output-skip setting (a request's last slice
      // keeps every column for its detector pass)
      for (int so = 0; so < 2; ++so) {
        std::vector<int> reqs;
        int btot = 0;
        for (int r : cls[c]) {
          if (s >= S[r]) continue;
          const bool last = s == S[r] - 1;
          if ((!rec_step && !last ? 1 : 0) != so) continue;
          reqs.push_back(r);
          btot += nprobe[r];
        }
        if (reqs.empty()) continue;
        const auto fc = add_chunks(reqs, one_row_chunk ? btot : row_pb(pc, btot), true, s);
        Op op{};
        op.kind = OP_ROW;
        op.first = fc.first;
        op.count = fc.second;
        op.mode = (s > 0 ? ROW_PRE_IFFT : 0u) | ROW_MUL_V | ROW_POST_FFT | ROW_CHUNKS;
        op.skip_in = s > 0;
        op.skip_out = so;
        ops[c].push_back(op);
      }
      // detector sums of the recorded slices
      for (int r : cls[c]) {
        if (s >= S[r] || !(rec_step || s == S[r] - 1)) continue;
        Op op{};
        op.kind = OP_DET;
        op.req = r;
        op.io = io[r]++;
        ops[c].push_back(op);
      }
      // propagator column pass of every request with slices left
      std::vector<int> reqs;
      int btot = 0;
      for (int r : cls[c])
        if (s < S[r] - 1) {
          reqs.push_back(r);
          btot += nprobe[r];
        }
      if (!reqs.empty()) {
        const int blocks = pc->grid.nx / col_tile(pc->grid.ny);
        const int target = std::max(1, pc->tune.col_blocks);
        // a class of at most two probes on such large grids: one chunk's
        // two-probe workgroups take as many rounds as two one-probe chunks
        // (a column workgroup holds a CU) and load P once (+1.6% on the mix)
        const int groups = (one_row_chunk && btot <= 2) ? 1
                                                        : std::max(1, std::min(btot, (target + blocks - 1) / blocks));
        const auto fc = add_chunks(reqs, (btot + groups - 1) / groups, false, s);
        Op op{};
        op.kind = OP_COL;
        op.first = fc.first;
        op.count = fc.second;
        ops[c].push_back(op);
      }
    }
  }
  for (const auto& oc : ops)
    for (const Op& op : oc)
      if (op.kind != OP_DET && op.count > 65535) return p0->fail(EM_ERR_INVALID, "too many probes in one pass");

  // size every buffer before capture (and before the key reads its address)
  for (int r = 0; r < nreq; ++r) {
    em_plan p = plans[r];
    const int ntiles = col_ntiles(p->grid.nx, p->grid.ny);
    EM_HIP(p, p->dpart.ensure(sizeof(float) * (size_t)nprobe[r] * ndet * ntiles));
  }

  // Everything a capture bakes in: the requests' buffers and slice counts, the
  // detector bands and stride, each plan's grid, wavelength, launch shapes,
  // skip ranges and tables, and the chunk table built from them.
  std::string key;
  auto put = [&key](const void* d, size_t n) { key.append(reinterpret_cast<const char*>(d), n); };
  put(&nreq, sizeof(nreq));
  put(&ndet, sizeof(ndet));
  put(&stride, sizeof(stride));
  put(inner, sizeof(double) * ndet);
  put(outer, sizeof(double) * ndet);
  for (int r = 0; r < nreq; ++r) {
    const em_plan p = plans[r];
    const void* ptrs[9] = {p, V[r], psi[r], out[r], p->dpart.p, p->P.p, p->tw_x.p, p->tw_y.p, p->twt_x.p};
    put(ptrs, sizeof(ptrs));
    put(&p->twt_y.p, sizeof(void*));
    put(&S[r], sizeof(int));
    put(&nprobe[r], sizeof(int));
    put(&p->tune, sizeof(em_tuning));
    const int sk[6] = {p->skip_lo, p->skip_hi, p->skip_lo_y, p->skip_hi_y, p->zk_x, p->zk_y};
    put(sk, sizeof(sk));
    put(&p->k.lambda, sizeof(double));
    const int gi[3] = {p->grid.nx, p->grid.ny, p->grid.bwl};
    const double gd[3] = {p->grid.lx, p->grid.ly, p->grid.dz};
    put(gi, sizeof(gi));
    put(gd, sizeof(gd));
  }
  put(chunks.data(), sizeof(PassChunk) * chunks.size());

  em_plan_s::RaggedGraph* rg = nullptr;
  for (auto& e : p0->ragged_graphs)
    if (e->key == key) rg = e.get();
  if (!rg) {
    auto entry = std::make_unique<em_plan_s::RaggedGraph>();
    entry->key = key;
    EM_HIP(p0, entry->chunks.ensure(sizeof(PassChunk) * chunks.size()));
    EM_HIP(p0, hipMemcpy(entry->chunks.p, chunks.data(), sizeof(PassChunk) * chunks.size(), hipMemcpyHostToDevice));
    const PassChunk* dchunks = entry->chunks.as<PassChunk>();

    EventSet evs;
    std::vector<hipEvent_t> done(cls.size(), nullptr);
    for (auto& e : done) EM_HIP(p0, evs.make(&e, hipEventDisableTiming));
    em_status st = p0->hip(hipStreamBeginCapture(p0->gstream, hipStreamCaptureModeThreadLocal), "begin capture");
    hipEvent_t cap_fork = nullptr;
    if (st == EM_OK) st = p0->hip(evs.make(&cap_fork, hipEventDisableTiming), "event");
    if (st == EM_OK) st = p0->hip(hipEventRecord(cap_fork, p0->gstream), "record");
    for (size_t c = 0; c < cls.size() && st == EM_OK; ++c) {
      // each class is a branch of the graph on its first plan's private stream
      // (class 0 holds request 0)
      em_plan pc = plans[cls[c][0]];
      hipStream_t br = pc->gstream;
      if (c > 0) st = pc->hip(hipStreamWaitEvent(br, cap_fork, 0), "fork");
      for (size_t k = 0; k < ops[c].size() && st == EM_OK; ++k) {
        const Op& op = ops[c][k];
        if (op.kind == OP_ROW) {
          RowArgs ra = row_args(pc, nullptr, nullptr, op.count);
          ra.pb = 1;
          ra.chunks = dchunks + op.first;
          ra.skip_in = op.skip_in;
          ra.skip_out = op.skip_out;
          ra.zk = (op.skip_in && op.skip_out) ? pc->zk_x : 0;
          st = pc->hip(launch_row(op.mode, ra, br), "row pass");
        } else if (op.kind == OP_COL) {
          ColArgs cc = col_args(pc, nullptr, nullptr, op.count);
          cc.pb = 1;
          cc.chunks = dchunks + op.first;
          cc.cols = 0;
          cc.variant = COL_VARIANT_BLOCK;
          cc.skip_lo = pc->skip_lo;
          cc.skip_hi = pc->skip_hi;
          cc.zk = pc->zk_y;
          st = pc->hip(launch_col(COL_PRE_FFT | COL_MUL_P | COL_POST_IFFT | COL_STORE | COL_CHUNKS, cc, br),
                       "column pass");
        } else {
          const int r = op.req;
          em_plan p = plans[r];
          const int B = nprobe[r];
          ColArgs cd = col_args(p, reinterpret_cast<float2*>(psi[r]), nullptr, B);
          cd.ndet = ndet;
          for (int d = 0; d < ndet; ++d) {
            cd.dg2min[d] = mrad_to_g2(inner[d], p->k.lambda, false);
            cd.dg2max[d] = mrad_to_g2(outer[d], p->k.lambda, true);
          }
          cd.dpart = p->dpart.as<float>();
          st = p->hip(launch_col(COL_PRE_FFT | COL_DET, cd, br), "column pass");
          if (st == EM_OK)
            st = p->hip(launch_det_reduce(B, ndet, col_ntiles(p->grid.nx, p->grid.ny), p->dpart.as<float>(), 1.f, 0,
                                          out[r] + (long long)op.io * ndet * B, B, br),
                        "detector reduce");
        }
      }
      if (st == EM_OK) st = pc->hip(hipEventRecord(done[c], br), "record");
      if (st == EM_OK && c > 0) st = pc->hip(hipStreamWaitEvent(p0->gstream, done[c], 0), "join");
    }
    const hipError_t ec = hipStreamEndCapture(p0->gstream, &entry->graph.g);
    if (st == EM_OK) st = p0->hip(ec, "end capture");
    if (st == EM_OK) st = p0->hip(hipGraphInstantiate(&entry->graph.x, entry->graph.g, nullptr, nullptr, 0), "instantiate");
    if (st != EM_OK) return st;
    if (p0->ragged_graphs.size() >= 4) {
      // the oldest graph may still run on the graph stream
      EM_HIP(p0, hipStreamSynchronize(p0->gstream));
      p0->ragged_graphs.erase(p0->ragged_graphs.begin());
    }
    rg = entry.get();
    p0->ragged_graphs.push_back(std::move(entry));
  }

  // the graph stream runs after the caller's stream, and the caller's stream
  // after it: no host synchronisation
  EM_HIP(p0, hipEventRecord(p0->rag_in, user));
  EM_HIP(p0, hipStreamWaitEvent(p0->gstream, p0->rag_in, 0));
  EM_HIP(p0, hipGraphLaunch(rg->graph.x, p0->gstream));
  EM_HIP(p0, hipEventRecord(p0->rag_out, p0->gstream));
  EM_HIP(p0, hipStreamWaitEvent(user, p0->rag_out, 0));
  return EM_OK;
}

}  // extern "C"

extern "C" em_status em_time_passes(em_plan p, const float* V, void* wave, int nprobe, int iters, float* row_ms,
                                    float* col_ms) {
  if (!p || !V || !wave || nprobe <= 0 || iters <= 0 || !row_ms || !col_ms) return EM_ERR_INVALID;
  float2* w = reinterpret_cast<float2*>(wave);
  EventSet evs;
  hipEvent_t e[3] = {nullptr, nullptr, nullptr};
  for (auto& x : e) EM_HIP(p, evs.make(&x));
  // the row pass as em_stem runs it: transmission precomputed per slice
  EM_HIP(p, p->tmp.ensure(sizeof(float2) * p->N()));
  EM_HIP(p, launch_transmission(V, p->tmp.as<float2>(), p->N(), p->stream));
  RowArgs r = row_args_skip(p, w, w, nprobe, true, true);
  r.t = p->tmp.as<float2>();
  const ColArgs c = col_args_skip(p, w, w, nprobe);
  em_status st = EM_OK;
  // warm up both kernels once
  unsigned rm = row_t_mode(p, ROW_PRE_IFFT | ROW_MUL_T | ROW_POST_FFT);
  if (p->tune.row_variant == 2) {
    r.V = V;
    rm = ROW_PRE_IFFT | ROW_MUL_V | ROW_POST_FFT;
  }
  st = row(p, rm, r);
  if (st == EM_OK) st = col(p, COL_PRE_FFT | COL_MUL_P | COL_POST_IFFT | COL_STORE, c);
  if (st == EM_OK) st = p->hip(hipEventRecord(e[0], p->stream), "event");
  for (int i = 0; i < iters && st == EM_OK; ++i) st = row(p, rm, r);
  if (st == EM_OK) st = p->hip(hipEventRecord(e[1], p->stream), "event");
  for (int i = 0; i < iters && st == EM_OK; ++i)
    st = col(p, COL_PRE_FFT | COL_MUL_P | COL_POST_IFFT | COL_STORE, c);
  if (st == EM_OK) st = p->hip(hipEventRecord(e[2], p->stream), "event");
  if (st == EM_OK) st = p->hip(hipEventSynchronize(e[2]), "event sync");
  float a = 0, b = 0;
  if (st == EM_OK) st = p->hip(hipEventElapsedTime(&a, e[0], e[1]), "elapsed");
  if (st == EM_OK) st = p->hip(hipEventElapsedTime(&b, e[1], e[2]), "elapsed");
  *row_ms = a / iters;
  *col_ms = b / iters;
  return st;
}

extern "C" em_status em_host_slicing(long long n, const double* z, double dz, int border, int max_slices,
                                     int* n_slices, double* bounds, long long* i0, long long* i1) {
  if (n < 0 || (n > 0 && !z) || !(dz > 0) || border < 0 || border > 2 || !n_slices) return EM_ERR_INVALID;
  for (long long i = 1; i < n; ++i)
    if (z[i] < z[i - 1]) return EM_ERR_INVALID;
  const std::vector<double> zz(z, z + n);
  const std::vector<SliceRange> s = make_slices(zz, dz, border, kZEps);
  *n_slices = (int)s.size();
  if ((int)s.size() > max_slices) return EM_ERR_INVALID;
  for (size_t k = 0; k < s.size(); ++k) {
    if (bounds) {
      bounds[k] = s[k].z_start;
      bounds[k + 1] = s[k].z_end;
    }
    if (i0) i0[k] = s[k].i0;
    if (i1) i1[k] = s[k].i1;
  }
  return EM_OK;
}

extern "C" em_status em_host_quadrature(int kind, int n, double* x, double* w) {
  if (n < 1 || n > 4096 || !x || !w || kind < 0 || kind > 1) return EM_ERR_INVALID;
  std::vector<double> xs, ws;
  if (kind == 0) gauss_legendre(n, xs, ws); else gauss_hermite(n, xs, ws);
  for (int i = 0; i < n; ++i) {
    x[i] = xs[i];
    w[i] = ws[i];
  }
  return EM_OK;
}

extern "C" em_status em_host_slicing_fine(long long n, const double* z, double dz, int border, double r_int,
                                          int max_slices, int* n_slices, double* bounds, long long* i0,
                                          long long* i1) {
  if (n < 0 || (n > 0 && !z) || !(dz > 0) || border < 0 || border > 2 || !(r_int >= 0) || !n_slices)
    return EM_ERR_INVALID;
  for (long long i = 1; i < n; ++i)
    if (z[i] < z[i - 1]) return EM_ERR_INVALID;
  const std::vector<double> zz(z, z + n);
  const std::vector<SliceRange> s = make_slices_fine(zz, dz, border, kZEps, r_int);
  *n_slices = (int)s.size();
  if ((int)s.size() > max_slices) return EM_ERR_INVALID;
  for (size_t k = 0; k < s.size(); ++k) {
    if (bounds) {
      bounds[k] = s[k].z_start;
      bounds[k + 1] = s[k].z_end;
    }
    if (i0) i0[k] = s[k].i0;
    if (i1) i1[k] = s[k].i1;
  }
  return EM_OK;
}

extern "C" em_status em_host_spline(int Z, int nR, double r_min, double r_max, double tap, double* R2, double* c0,
                                    double* c1, double* c2, double* c3) {
  if (Z <= 0 || Z > 103 || nR < 4 || nR > 256 || !(r_min > 0) || !(r_max > r_min) || !(tap > 0) || !(tap < 1))
    return EM_ERR_INVALID;
  const Spline s = build_spline(moliere_coefs(Z), nR, r_min, r_max, tap);
  for (int i = 0; i < nR; ++i) {
    if (R2) R2[i] = s.R2[i];
    if (c0) c0[i] = s.c0[i];
    if (c1) c1[i] = s.c1[i];
    if (c2) c2[i] = s.c2[i];
    if (c3) c3[i] = s.c3[i];
  }
  return EM_OK;
}

extern "C" em_status em_host_constants(double e0_kv, double* lambda, double* gamma, double* sigma) {
  if (!(e0_kv > 0)) return EM_ERR_INVALID;
  const Constants k = beam_constants(e0_kv);
  if (lambda) *lambda = k.lambda;
  if (gamma) *gamma = k.gamma;
  if (sigma) *sigma = k.sigma;
  return EM_OK;
}

constexpr double kTuneMargin = 0.02;

extern "C" em_status em_plan_autotune(em_plan p, int nprobe, int nslices, double* best) {
  if (!p || nprobe <= 0 || nslices <= 0) return EM_ERR_INVALID;
  const long long N = p->N();
  // timed on precomputed transmissions, the form the STEM slice loop runs
  // (on potential slices for row_variant 2)
  float2* T = nullptr;
  float2* w = nullptr;
  EM_HIP(p, hipMalloc(&T, sizeof(float2) * N * nslices));
  if (hipMalloc(&w, sizeof(float2) * N * nprobe) != hipSuccess) {
    (void)hipFree(T);
    return p->fail(EM_ERR_ALLOC, "autotune scratch");
  }
  // potential slices: the real parts of the unit phasors (any values do)
  float* Vt = reinterpret_cast<float*>(T);
  EventSet evs;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  em_status st = p->hip(launch_phase_fill(T, N * nslices, p->stream), "phase fill");
  if (st == EM_OK) st = p->hip(launch_phase_fill(w, N * nprobe, p->stream), "phase fill");
  if (st == EM_OK) st = p->hip(evs.make(&e0), "event");
  if (st == EM_OK) st = p->hip(evs.make(&e1), "event");
  // best of three timed runs after a warm-up run: the minimum is the most
  // repeatable statistic when other work on the device adds noise
  auto measure = [&](double* ms) -> em_status {
    const bool vmode = p->tune.row_variant == 2;
    const float* Vm = vmode ? Vt : nullptr;
    const void* Tm = vmode ? nullptr : T;
    EM_TRY(em_multislice_forward(p, Vm, Tm, nslices, w, nprobe, nullptr, 1));  // warm
    float best = 0;
    for (int i = 0; i < 3; ++i) {
      EM_HIP(p, hipEventRecord(e0, p->stream));
      EM_TRY(em_multislice_forward(p, Vm, Tm, nslices, w, nprobe, nullptr, 1));
      EM_HIP(p, hipEventRecord(e1, p->stream));
      EM_HIP(p, hipEventSynchronize(e1));
      float t = 0;
      EM_HIP(p, hipEventElapsedTime(&t, e0, e1));
      if (i == 0 || t < best) best = t;
    }
    *ms = best / ((double)nprobe * nslices);
    return EM_OK;
  };
  double cur = 0;
  const em_tuning start = p->tune;
  if (st == EM_OK) st = measure(&cur);
  // coordinate descent over the launch-shape parameters, two sweeps
  const int groups[] = {0, 8, 16, 24, 32, 40, 48, 64, nprobe};
  // row-pass targets: powers of two, and whole rounds of the workgroups the
  // chip holds at once (1024^2: 256 CUs x 3 = 768 ran 3% faster per
  // probe-slice than 2048, whose last round is partly empty)
  std::vector<int> rowb = {512, 1024, 2048, 4096, 8192, 16384, whole_round_row_target(p)};
  if (const int slots = row_slots(p))
    for (int k : {1, 2, 4}) rowb.push_back(k * slots);
  std::sort(rowb.begin(), rowb.end());
  rowb.erase(std::unique(rowb.begin(), rowb.end()), rowb.end());
  const int colb[] = {256, 512, 1024, 2048, 4096, 8192, 16384};
  const int tiles[] = {0, col_alt_tile(p->grid.ny)};
  const int variants[] = {1, 2};
  const int row_variants[] = {0, 1, 2};
  // the persistent slice loop where the grid supports it
  const int persistents[] = {0, 1};
  const int npers = persist_supported(p->grid.nx, p->grid.ny, p->zk_x, p->zk_y) ? 2 : 1;
  // radix sequences of the row and column lines
  const int radices[] = {0, 1, 2, 3};
  const int nrx = fft_variants(p->grid.nx), nry = fft_variants(p->grid.ny);
  for (int sweep = 0; sweep < 2 && st == EM_OK; ++sweep) {
    for (int par = 0; par < 9 && st == EM_OK; ++par) {
      const int* cand = par == 0   ? groups
                        : par == 1 ? rowb.data()
                        : par == 2 ? colb
                        : par == 3 ? tiles
                        : par == 4 ? variants
                        : par == 5 ? row_variants
                        : par == 6 ? persistents
                                   : radices;
      const int nc = par == 0   ? 9
                     : par == 1 ? (int)rowb.size()
                     : par == 2 ? 7
                     : par == 5 ? 3
                     : par == 6 ? npers
                     : par == 7 ? nrx
                     : par == 8 ? nry
                                : 2;
      em_tuning bestt = p->tune;
      for (int i = 0; i < nc && st == EM_OK; ++i) {
        em_tuning t = p->tune;
        if (par == 0) t.probe_group = cand[i];
        if (par == 1) t.row_blocks = cand[i];
        if (par == 2) t.col_blocks = cand[i];
        if (par == 3) {
          if (cand[i] != 0 && p->grid.nx % cand[i] != 0) continue;
          t.col_tile = cand[i];
        }
        if (par == 4) t.col_variant = cand[i];
        if (par == 5) t.row_variant = cand[i];
        if (par == 6) t.persistent = cand[i];
        if (par == 7) t.row_radix = cand[i];
        if (par == 8) t.col_radix = cand[i];
        if (par == 0 && cand[i] > nprobe) continue;
        const em_tuning keep = p->tune;
        p->tune = t;
        double ms = 0;
        st = measure(&ms);
        if (st == EM_OK && ms < cur) {
          cur = ms;
          bestt = t;
        }
        p->tune = keep;
      }
      p->tune = bestt;
    }
  }
  // The descent keeps the fastest of many noisy trials, which favours lucky
  // ones: re-time the choice against the starting shapes, alternately, and
  // keep it only if it is at least kTuneMargin faster.  Run-to-run noise is
  // about 1%; without the margin, boxes whose best shapes tie with the
  // defaults flip between them from run to run.
  const em_tuning chosen = p->tune;
  if (st == EM_OK && std::memcmp(&chosen, &start, sizeof(em_tuning)) != 0) {
    double a = 0, b = 0;
    for (int r = 0; r < 3 && st == EM_OK; ++r) {
      double ms = 0;
      p->tune = start;
      st = measure(&ms);
      if (st != EM_OK) break;
      a = (r == 0) ? ms : std::min(a, ms);
      p->tune = chosen;
      st = measure(&ms);
      b = (r == 0) ? ms : std::min(b, ms);
    }
    const bool take = st != EM_OK || b <= a * (1.0 - kTuneMargin);
    p->tune = take ? chosen : start;
    cur = take ? b : a;
  }
  (void)hipStreamSynchronize(p->stream);
  (void)hipFree(T);
  (void)hipFree(w);
  if (best) *best = cur;
  return st;
}
