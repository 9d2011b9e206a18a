// em_kernels.hpp -- host-side declarations of the gfx950 kernels' launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace em {

// Grid description passed by value to the kernels.
struct GridDev {
  int nx, ny;
  float dx, dy;      // real-space sampling (Angstrom)
  double dgx, dgy;   // reciprocal sampling (1/Angstrom), kept in FP64 so band
                     // membership of a pixel is decided exactly as on the host
  double gl2_max, alpha;  // band-width-limit Fermi factor (grid_dev.cuh:53-55)
};

// Lens in derived form (lens_dev.cuh:40-57), FP64: the probe and CTF tables
// are built once per run, so the aberration phase is evaluated in double.
struct LensDev {
  double m, cc10, cc12, cc21, cc23, cc30, cc32, cc34, cc41, cc43, cc45, cc50, cc52, cc54, cc56;
  double phi12, phi21, phi23, phi32, phi34, phi41, phi43, phi45, phi52, phi54, phi56;
  double g2_min, g2_max;
  int phi_required;
  // tilt added to g (apply_CTF_kernel's gxu, gyu, device_kernels.cuh:446-454)
  double gxs = 0, gys = 0;
  // quasi-coherent envelope of the PCTF (incoherent_phase_shift,
  // basic_functions_dev.cuh:236-266): on/off, wavelength, defocus and C30 (A),
  // defocus-spread standard deviation (A), source half-width angle (rad)
  int envelope = 0;
  double env_lambda = 0, env_c10 = 0, env_c30 = 0, env_sigma = 0, env_beta = 0;
};

// One entry of a chunk table (ragged batches): n probes of one request,
// pstride elements apart, read from `in` and written to `out`, with that
// request's own table -- the potential slice V (row pass) or the propagator
// P (column pass).  The passes run in place (in == out); the two fields keep
// the kernels from holding every load address for the matching store.
struct PassChunk {
  const float2* in;
  float2* out;
  const void* table;
  int n;
  int pad;
};

// --- row pass ---------------------------------------------------------------
enum RowMode : unsigned {
  ROW_PRE_IFFT = 1u << 0,    // inverse FFT along x before the multiply
  ROW_POST_FFT = 1u << 1,    // forward FFT along x after the multiply
  ROW_MUL_V = 1u << 2,       // multiply by exp(i V), V sigma-scaled FP32
  ROW_MUL_T = 1u << 3,       // multiply by a complex transmission t
  ROW_CONJ = 1u << 4,        // use conj of the transmission (adjoint)
  ROW_IN_PLANE = 1u << 5,    // input is the plane wave 1 (no load)
  ROW_CKPT_STORE = 1u << 6,  // store the real-space wave before the multiply
  ROW_GRAD = 1u << 7,        // accumulate conj(ckpt) * wave before the multiply
  ROW_OUT_ABS2 = 1u << 8,    // store |.|^2 (FP32) instead of the complex wave
  ROW_OUT_SPLIT = 1u << 9,   // store re / im planes
  ROW_OUT_ACCUM = 1u << 10,  // accumulate weight * output into the FP32 planes
  ROW_T_STREAM = 1u << 11,   // read t for every probe instead of holding it in
                             // registers (frees 2E VGPRs for a fourth wave per SIMD)
  ROW_CHUNKS = 1u << 12,     // in place; workgroup row y takes its probes, wave and
                             // V from chunks[y] (launch with nprobe = chunks, pb = 1)
};

struct RowArgs {
  const float2* in;
  float2* out;
  long long pstride;        // elements between probes
  int nx, ny, nprobe, pb;   // pb: probes per workgroup
  const float* V;           // [ny][nx]
  const float2* t;          // [ny][nx]
  const float2* tw;         // [nx] exp(-2 pi i m / nx)
  const float2* twt;        // radix variant 3: the per-lane twiddle table (tangent_tables)
  float2* ckpt;             // [nprobe][ny][nx]
  float2* gpart;            // [gridDim.y][ny][nx] gradient partials
  float* o1;                // [nprobe][ny][nx] intensity or real part
  float* o2;                // [nprobe][ny][nx] imaginary part
  float wgt;
  // Band-limit skipping: columns kx in [skip_lo, skip_hi] carry nothing the
  // next column pass keeps (the propagator is below the skip threshold there).
  // skip_in: those input columns are taken as zero and not read;
  // skip_out: those output columns are not written.
  int skip_lo, skip_hi, skip_in, skip_out;
  // slots per line end wholly inside [skip_lo, skip_hi] (pruned_slots): with
  // skip_in and skip_out set, the slice-loop modes then run pruned
  int zk;
  int radix;                // radix sequence of the pruned STEM row pass (fft_variants)
  const PassChunk* chunks;  // ROW_CHUNKS: one entry per workgroup row
};

// --- column pass ------------------------------------------------------------
enum ColMode : unsigned {
  COL_PRE_FFT = 1u << 0,
  COL_PRE_IFFT = 1u << 1,
  COL_MUL_P = 1u << 2,        // multiply by the propagator
  COL_CONJ_P = 1u << 3,       // ... by its conjugate (adjoint)
  COL_MUL_FILTER = 1u << 4,   // closed band [g2min, g2max] with 1/N (Filter_wave)
  COL_POST_FFT = 1u << 5,
  COL_POST_IFFT = 1u << 6,
  COL_LOAD_PROBE = 1u << 7,   // value = A * ex[b][x] * ey[b][y]; nothing loaded
  COL_LOAD_MUL_A = 1u << 8,   // value = wave * A
  COL_STORE = 1u << 9,
  COL_DET = 1u << 10,         // detector partial sums over closed bands
  COL_CHUNKS = 1u << 11,      // in place; workgroup row y takes its probes, wave and
                              // P from chunks[y] (launch with nprobe = chunks, pb = 1;
                              // default tile, block variant)
};

constexpr int kMaxDet = 8;

struct ColArgs {
  const float2* in;
  float2* out;
  long long pstride;
  int nx, ny, nprobe, pb;
  const float2* P;          // [ny][nx]
  const float2* tw;         // [ny]
  const float2* twt;        // radix variant 3: the per-lane twiddle table (tangent_tables)
  const float2* A;          // [ny][nx]
  const float2* ex;         // [nprobe][nx]
  const float2* ey;         // [nprobe][ny]
  GridDev grid;
  double fg2min, fg2max;    // filter band
  float fscale;
  int ndet;
  double dg2min[kMaxDet], dg2max[kMaxDet];
  float* dpart;             // [nprobe][ndet][ntiles]
  int cols;                 // column tile of the per-slice modes: 0 = default, or col_alt_tile
  int variant;              // per-slice modes: COL_VARIANT_BLOCK or COL_VARIANT_TRANSPOSED
  // per-slice modes only: column tiles wholly inside [skip_lo, skip_hi] are
  // neither read nor written (set from skip_lo/hi by the launcher)
  int skip_lo, skip_hi;
  int tile_lo, tile_hi;
  // propagator modes: ky slots wholly inside the band-limit skip range of
  // the column (pruned_slots of ny and that range); > 0 runs them pruned
  int zk;
  int radix;                // radix sequence of the pruned propagator pass (fft_variants)
  const PassChunk* chunks;  // COL_CHUNKS: one entry per workgroup row
};

enum { COL_VARIANT_BLOCK = 0, COL_VARIANT_TRANSPOSED = 1 };

// --- persistent slice loop -----------------------------------------------------
// One launch over a probe group: the propagator column passes of slices
// 0..S-2 and the row passes of slices 1..S-2 between them (all pruned; the
// caller runs slice 0's row pass before and slice S-1's after).  S >= 2.
// ctr: persist_counters(S) ints, zero before the
// launch (per-XCD work queues of every phase, the barrier count); *err is set
// to 1 when a barrier wait passes spin_limit (every workgroup then leaves).
struct PersistArgs {
  float2* wave;             // [nprobe] waves, pstride apart, (y, kx) layout
  long long pstride;
  int n, nprobe, S;
  const float2* t;          // [S][n][n] transmissions
  const float2* P;          // [n][n]
  const float2* twx;
  const float2* twy;
  int skip_lo, skip_hi;     // band-limit skip (row passes)
  int tile_lo, tile_hi;     // 8-column tiles wholly inside it (column passes)
  int zk_x, zk_y;           // line slots the pruned items may drop
  int row_pb, col_pb;       // probes per work item
  int* ctr;
  int* err;
  long long spin_limit;     // wall-clock ticks (100 MHz) of one barrier wait
};
bool persist_supported(int nx, int ny, int zk_x, int zk_y);
int persist_counters(int S);
// workgroups of a launch (all resident: occupancy x CUs, a multiple of 8)
hipError_t persist_grid(int n, int* grid);
hipError_t launch_persist(const PersistArgs& a, int grid, hipStream_t s);

hipError_t launch_row(unsigned mode, const RowArgs& a, hipStream_t s);
hipError_t launch_col(unsigned mode, const ColArgs& a, hipStream_t s);
// Row-pass workgroups of this mode and row length that one CU holds at once
// (the occupancy limit of its registers and LDS).
hipError_t row_blocks_per_cu(int nx, unsigned mode, int* n);
// Arms a pair of timing events for the next row or column pass launched from
// the calling thread: that launch records the kernel's own start and end in
// them (hipExtLaunchKernel), without the dispatch gap that events queued
// around a launch also measure.  pass_events_pending: armed and not yet taken.
void set_pass_events(hipEvent_t start, hipEvent_t stop);
bool pass_events_pending();
// Raises the dynamic-LDS limit of every column-pass kernel of length ny
// (call once per device before launching them).
hipError_t prepare_col_kernels(int ny);
bool fft_size_supported(int n);
// radix sequences a line length is built with (tuning row_radix / col_radix)
int fft_variants(int n);
int row_lines_per_block(int nx);
int col_tile(int ny);      // default columns per column-pass workgroup
int col_alt_tile(int ny);  // the alternative tile of the per-slice column modes
// whether the transposed-exchange column pass exists for this column length
bool col_transposed_available(int ny);
int col_ntiles(int nx, int ny);
// Slots per line end that a band-limit skip range [lo, hi] of a line of n
// points covers wholly (the pruned passes need at least the slots they are
// built for).
int pruned_slots(int n, int lo, int hi);

// --- other kernels ------------------------------------------------------------
// fine: null, or one float4 (c0, c1, c2, c3) per spline node of every atom
// entry (fine slicing: a spline per atom and slice instead of per element)
hipError_t launch_potential(const GridDev& g, int nslices, const int* slice_bin_off,
                            int nbx, int nby, float bin_size, const float4* atoms,
                            int ntypes, int nR, const float* tables, const float* rmax2,
                            const float4* fine, float* V, hipStream_t s);

// 3D potential of one element for fine slicing: parameterisation type pt
// (1 Doyle-Turner 0-4, 2 Peng 0-4, 3 Peng 0-12, 4 Kirkland 0-12,
// 5 Weickenmeier-Kohl 0-12, 6 Lobato 0-12), coefficients, R^2 taper.
struct FineElement {
  int pt;
  double cl[6], cnl[6];
  double r2_tap, tap_cf;
};
constexpr int kMaxFineQ = 1024;
// Spline of every atom entry e (type in atoms[e].w, slice centre height above
// the atom az[e]) over its slice of thickness dz: nR nodes R2[type][i]
// (double), nq Gauss-Legendre points (x, w) on [-1/2, 1/2] -> fine[e][nR].
hipError_t launch_fine_coefs(long long nentries, int nR, const double* R2, const FineElement* el,
                             const float4* atoms, const float* az, double dz, int nq,
                             const double2* q, float4* fine, hipStream_t s);
hipError_t launch_propagator(const GridDev& g, double w, double gx0, double gy0, int bwl,
                             float2* P, hipStream_t s);
hipError_t launch_lens_table(const GridDev& g, const LensDev& l, float2* A, hipStream_t s);
hipError_t launch_probe_ramps(const GridDev& g, int nprobe, const double* xe, const double* ye,
                              float2* ex, float2* ey, hipStream_t s);
// out[d * ostride + b] (+)= w * sum_tiles part[b][d][tile]
hipError_t launch_det_reduce(int nprobe, int ndet, int ntiles, const float* part, float w,
                             int accumulate, float* out, long long ostride, hipStream_t s);
// g (+)= sum_p part[p]  (accumulate: add to g instead of overwriting it)
hipError_t launch_grad_reduce(long long n, int nparts, const float2* part, float2* g, int accumulate,
                              hipStream_t s);
hipError_t launch_fill(float2* p, long long n, float2 v, hipStream_t s);
hipError_t launch_phase_fill(float2* p, long long n, hipStream_t s);
hipError_t launch_fft_shift(const float* in, float* out, int nx, int ny, long long nbatch, float w,
                            int accumulate, hipStream_t s);
hipError_t launch_abs2(const float2* in, float* out, long long n, float w, int accumulate,
                       hipStream_t s);
hipError_t launch_split(const float2* in, float* re, float* im, long long n, float w,
                        int accumulate, hipStream_t s);
hipError_t launch_transmission(const float* V, float2* t, long long n, hipStream_t s);
hipError_t launch_grad_v(const float2* gt, const float* V, float sigma, long long n, float* gV,
                         hipStream_t s);

}  // namespace em
