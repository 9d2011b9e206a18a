/* cudaem_amd.h -- C ABI of the MI355X (gfx950) multislice engine.
 *
 * The reference (ningustc/cudaEM) exposes header-only C++ templates and no
 * C ABI; its device work happens inside Wave_Function, Transmission_Function,
 * Propagator, Incident_Wave and Lens_Modulation.  Each entry point below names
 * the reference code it replaces (file:line, files of the reference's flat
 * directory).  The C++ simulator surface (include/cudaEM/simulator.hpp) and
 * the Python package (cudaem_amd/) are built on these entry points.
 *
 * Conventions
 *  - Every function returns an em_status; nothing throws across the ABI.
 *  - Pointers named *_dev are device pointers owned by the caller; all other
 *    pointers are host pointers.  Complex device data is interleaved FP32
 *    (float2), row-major [row][col] with col (x) fastest, batches outermost.
 *  - Work is enqueued on the plan's stream (default: the null stream) and is
 *    asynchronous unless the function returns host data.
 *  - One plan per host thread or stream; a plan is not thread-safe.
 *  - FFT convention: forward exp(-2 pi i nk/N), inverse exp(+2 pi i nk/N),
 *    both unnormalised (fft_dev.cuh:159-190).  Exit waves are left in Fourier
 *    space, unnormalised (wave_function.cuh:60-73).
 */
#ifndef CUDAEM_AMD_H
#define CUDAEM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef int em_status;
#define EM_OK 0
#define EM_ERR_INVALID 1     /* bad argument or call order */
#define EM_ERR_UNSUPPORTED 2 /* grid size / mode not built for gfx950 */
#define EM_ERR_ALLOC 3       /* device or host allocation failed */
#define EM_ERR_HIP 4         /* a HIP runtime call failed */

typedef struct em_plan_s* em_plan;

/* Simulation box: Grid_2d_Dev (grid_dev.cuh:10-56). */
typedef struct em_grid {
  int nx, ny;       /* columns (col) and rows (row); each of 64..2048 from the built set */
  double lx, ly;    /* field of view, Angstrom (x_l, y_l) */
  double dz;        /* slice thickness, Angstrom (dz) */
  int bwl;          /* 1: 2/3 band-width limit in the propagator (bwl) */
} em_grid;

/* Beam: acceleration voltage and tilt (multislice_params_dev.cuh:40-41, 201-208). */
typedef struct em_beam {
  double e0_kv;
  double theta, phi; /* rad */
} em_beam;

/* Lens: Lens_Dev (lens_dev.cuh:9-66), coefficients in Angstrom, angles in rad,
 * apertures as semi-angles in mrad (g = sin(angle)/lambda; outer <= 0: none). */
typedef struct em_lens {
  double c_10, c_12, phi_12, c_21, phi_21, c_23, phi_23, c_30, c_32, phi_32, c_34, phi_34,
      c_41, phi_41, c_43, phi_43, c_45, phi_45, c_50, c_52, phi_52, c_54, phi_54, c_56, phi_56;
  double m;
  double inner_aper_mrad, outer_aper_mrad;
} em_lens;

/* Atoms: Atom_Coord_Data (x, y, z, Z, occ); occ may be NULL (all 1). */
typedef struct em_atoms {
  long long n;
  const double* x;
  const double* y;
  const double* z;
  const int* Z;
  const double* occ;
} em_atoms;

/* Slice alignment: eMatch_Border (slice_generator.hpp:137-166). */
#define EM_BORDER_MIN 0    /* STEM / ED / CBED (specimen.cuh:33-36) */
#define EM_BORDER_MAX 1    /* HRTEM (specimen.cuh:29-32) */
#define EM_BORDER_MIDDLE 2

/* ---- plan ------------------------------------------------------------------
 * Replaces multislice_params_dev's device state and cuFFT plans
 * (multislice_params_dev.cuh:84-199) and Propagator::update_prop_function
 * (propagation_function.cuh:49-57, built once by wave_function.cuh:42-44).
 * max_batch bounds the probes of one call to the batched entry points. */
em_status em_plan_create(const em_grid* grid, const em_beam* beam, int max_batch, em_plan* out);
em_status em_plan_destroy(em_plan plan);
em_status em_plan_set_stream(em_plan plan, void* hip_stream);
/* lambda (A), sigma (rad/(V A)), gamma; Vr_factor = sigma / cos(theta)
 * (transmission_function.cuh:61-62). */
em_status em_plan_constants(em_plan plan, double* lambda, double* sigma, double* gamma);
/* Group of probes a fused pass keeps together (0 = automatic). */
em_status em_plan_set_probe_group(em_plan plan, int probes_per_group);
/* Probes per group the plan uses for a batch of nprobe probes: the target
 * group size, evened out over the batch's groups (the last may be smaller). */
em_status em_plan_group_size(em_plan plan, int nprobe, int* probes_per_group);
/* Per-launch timing of em_stem's slice loop (no reference counterpart: it
 * replaces reading kernel durations off a profiler).  While on, em_stem runs
 * its batches without graph capture, and every interior row pass and column
 * pass launch records its own start and end in a pair of HIP events
 * (hipExtLaunchKernel); em_plan_pass_times then returns the
 * average launch duration of each (ms) and the number of launches averaged,
 * and clears the records. */
em_status em_plan_set_pass_timing(em_plan plan, int on);
em_status em_plan_pass_times(em_plan plan, double* row_ms, double* col_ms, int* n_row, int* n_col);
/* Launch shape of the fused passes (no reference counterpart: the reference
 * relies on cuFFT's planner).  row_blocks / col_blocks: target workgroups of a
 * row / column pass (the probes each workgroup processes follow from it; the
 * row default is the fewest probe chunks whose workgroups fill whole rounds of
 * the row workgroups the device holds at once, e.g. 768 at 1024^2 on 256 CUs);
 * col_tile: columns per column-pass workgroup (0 = default, or the
 * alternative tile of the grid size); probe_group: probes carried together
 * through all slices (0 = about 320 MiB of waves, at least 64 waves of more
 * than 16 MiB, evened out over a batch: with band skipping the passes touch
 * about 3/4 of it, near the 256 MiB last-level cache); col_variant: exchange scheme of the per-slice
 * column pass (0 = automatic, 1 = workgroup barrier per exchange,
 * 2 = one wavefront per column after the first stage); row_variant: row
 * pass of the STEM slice loop (0 = precomputed transmission held in
 * registers, 1 = read per probe, four waves per SIMD, 2 = potential slice
 * read and its sincos held in registers: half the table traffic);
 * persistent: 1 = run the interior of a probe group's slice loop on
 * precomputed transmissions (the propagator passes of slices 0..S-2 and the
 * row passes between them) as one cooperative launch with grid barriers
 * (1024^2 grids with the band-limit skip on; the call then waits for its
 * work to finish, and reports a barrier wait that timed out as EM_ERR_HIP);
 * row_radix / col_radix: radix sequence of the interior STEM row pass and
 * of the propagator column pass (at 1024: 0 = 16.16.4, 1 = 16.4.16 -- a new
 * plan's choice for both passes --, 2 = 4.16.16, and 3 = rows 16.16.4 /
 * columns 16.4.16 with their per-lane twiddles in factored FMA form; at
 * 2048: 0 = 16.16.8, 1 = 16.8.16 -- a new plan's rows --, 2 = 8.16.16;
 * other sizes have only 0). */
typedef struct em_tuning {
  int row_blocks, col_blocks, col_tile, probe_group, col_variant, row_variant, persistent;
  int row_radix, col_radix;
} em_tuning;
em_status em_plan_get_tuning(em_plan plan, em_tuning* tuning);
em_status em_plan_set_tuning(em_plan plan, const em_tuning* tuning);
/* Autotuner: times candidate tunings (coordinate descent, two sweeps) by
 * running nslices slices of nprobe synthetic probes on the plan's grid with
 * HIP events, then re-times the fastest against the plan's starting tuning
 * and keeps it only if it is at least 2% faster; best_ms_per_probe_slice
 * receives the kept tuning's time.  Allocates and frees its own scratch. */
em_status em_plan_autotune(em_plan plan, int nprobe, int nslices, double* best_ms_per_probe_slice);
/* Band-limit skipping (no reference counterpart).  With the band-width limit
 * on, the propagator of every column kx with bwl factor <= threshold at ky = 0
 * is below threshold / N in every row, so the column passes of the slice loop
 * skip those columns and the row passes neither read nor write them (except
 * where the full exit wave is formed).  The skipped contributions are below
 * threshold relative; default 1e-9 (under FP32 resolution); 0 disables;
 * at most 1e-6.  skip_lo / skip_hi receive the skipped column range. */
em_status em_plan_set_band_skip(em_plan plan, double threshold, int* skip_lo, int* skip_hi);
/* Last error message of this plan (static storage, valid until the next call). */
const char* em_plan_last_error(em_plan plan);
int em_fft_size_supported(int n);

/* ---- specimen and potential ----------------------------------------------
 * Element tables: the reference loads them in the missing element_coeffs_loader.hpp.
 * Default: Kirkland's projected-potential form with Moliere coefficients per Z.
 * em_set_element_coefs replaces them for one Z with Kirkland coefficients
 * (a_i, b_i for the K0 terms, c_i, d_i for the Gaussian terms). */
em_status em_set_element_coefs(em_plan plan, int Z, int nk, const double* a, const double* b,
                               int ng, const double* c, const double* d);
/* Spline nodes: nR nodes log-spaced in R over [r_min, r_max], cosine taper
 * beyond tap_frac * r_max (defaults 128, 0.005, 3.0, 0.85). */
em_status em_set_spline_params(em_plan plan, int nR, double r_min, double r_max, double tap_frac);
/* Partial coherence of the objective lens (Lens_Dev dsf / ssf fields,
 * lens_dev.cuh:27-34).  dsf_sigma: standard deviation of the Gaussian defocus
 * spread (A); ssf_beta: e^-1 half-width semi-angle of the Gaussian source
 * (mrad), i.e. tilts q ~ exp(-|q|^2 / q0^2) with q0 = sin(beta) / lambda;
 * *_npoints: Gauss-Hermite points of the TCC integrals (defocus; per tilt
 * axis); mode: which spreads apply (eTemporal_Spatial_Incoh). */
enum { EM_COH_TEMPORAL_SPATIAL = 1, EM_COH_TEMPORAL = 2, EM_COH_SPATIAL = 3 };
typedef struct em_coherence {
  double dsf_sigma;
  int dsf_npoints;
  double ssf_beta;
  int ssf_npoints;
  int mode;
} em_coherence;

/* 3D potential parameterisations for fine slicing (Vr_dVrir_*,
 * basic_functions_dev.cuh:104-214; the reference's ePotential_Type values are
 * in a missing header, these numbers are the build's). */
enum {
  EM_PT_DOYLE_0_4 = 1,         /* 4 Gaussians cl exp(-cnl r^2) */
  EM_PT_PENG_0_4 = 2,          /* 5 Gaussians */
  EM_PT_PENG_0_12 = 3,         /* 5 Gaussians */
  EM_PT_KIRKLAND_0_12 = 4,     /* 3 Yukawa cl exp(-cnl r)/r + 3 Gaussians */
  EM_PT_WEICKENMEIER_0_12 = 5, /* 6 terms cl erfc(cnl r)/r */
  EM_PT_LOBATO_0_12 = 6        /* 5 terms cl (2/(cnl r) + 1) exp(-cnl r) */
};
/* The 3D potential of element Z for fine slicing: type and 6 linear (cl, V A
 * or V) and 6 non-linear (cnl, 1/A or 1/A^2) coefficients, unused terms
 * ignored (the per-element Vr_coeff of device_containers.cuh:239, 280).
 * Default for elements without one: Moliere's screening in the Kirkland form
 * (3 Yukawa terms), whose projection is the default projected potential. */
em_status em_set_element_params(em_plan plan, int Z, int ptype, const double* cl, const double* cnl);
/* Fine slicing on / off (multislice_params.enable_fine_slicing,
 * projected_potential.cuh:33-44, 80-88) with nq Gauss-Legendre points per
 * slice (default 128, at most 1024).  Applies from the next
 * em_plan_set_specimen: slices then span [z_min - R_max, z_max + R_max], each
 * holding the atoms within R_max of its centre, and every (atom, slice) pair
 * gets the spline of the atom's 3D potential integrated over that slice
 * (gen_sa_coeff_kernel + fill_spl_coeff_kernel, device_kernels.cuh:48-105). */
em_status em_plan_set_fine_slicing(em_plan plan, int enable, int nq);
/* Copies the atoms, sorts them by z, slices them (Specimen::set_input_data,
 * specimen.cuh:16-36 + Slice_Generator::calculate, slice_generator.hpp:48-168)
 * and uploads the binned atom list.  Returns the slice count. */
em_status em_plan_set_specimen(em_plan plan, const em_atoms* atoms, int border, int* n_slices);
em_status em_slice_info(em_plan plan, int islice, double* z_start, double* z_end,
                        long long* iatom_start, long long* iatom_end);
/* Sigma-scaled projected potential of slices [s0, s0+ns) into V_dev[ns][ny][nx]:
 * Projected_Potential::generate_potential_by_slice (projected_potential.cuh:47-92)
 * with gen_slice_potential_kernel (device_kernels.cuh:28-47) and coeff = Vr_factor. */
em_status em_potential_slices(em_plan plan, int s0, int ns, float* V_dev);
/* Spline tables actually used (host copies, type order = sorted distinct Z):
 * R2, c0..c3 each [ntypes][nR]; returns ntypes and nR. */
em_status em_spline_tables(em_plan plan, int max_types, int* Zs, int* ntypes, int* nR,
                           double* R2, double* c0, double* c1, double* c2, double* c3);

/* Host-only helpers (no device work; usable without a GPU). */
/* Slicing of z-sorted atom heights: Slice_Generator::gene_bounds / get_slicing /
 * fd_by_z (slice_generator.hpp:81-168).  bounds[S+1], iatom_start[S],
 * iatom_end[S] (inclusive; start > end for an empty slice); capacity = max_slices.
 * Returns EM_ERR_INVALID when more than max_slices slices would be made. */
em_status em_host_slicing(long long n, const double* z_sorted, double dz, int border, int max_slices,
                          int* n_slices, double* bounds, long long* iatom_start, long long* iatom_end);
/* Quadrature nodes the engine uses, host only: kind 0 = Gauss-Legendre on
 * [-1/2, 1/2] (fine slicing), kind 1 = Gauss-Hermite for exp(-t^2) (TCC). */
em_status em_host_quadrature(int kind, int n, double* x, double* w);
/* Fine-sliced slicing (slice_generator.hpp:88-93, 108-128), host only: as
 * em_host_slicing over [z_min - r_int, z_max + r_int] with atom windows
 * [min(z_mid - r_int, z_start) - eps, max(z_mid + r_int, z_end) + eps]. */
em_status em_host_slicing_fine(long long n, const double* z_sorted, double dz, int border, double r_int,
                               int max_slices, int* n_slices, double* bounds, long long* iatom_start,
                               long long* iatom_end);
/* Spline of the projected potential of element Z (default coefficients) in
 * R^2: R2, c0..c3 each [nR] (basic_functions_dev.cuh:68-101 conventions). */
em_status em_host_spline(int Z, int nR, double r_min, double r_max, double tap_frac, double* R2,
                         double* c0, double* c1, double* c2, double* c3);
/* lambda, gamma, sigma for an acceleration voltage (kV). */
em_status em_host_constants(double e0_kv, double* lambda, double* gamma, double* sigma);

/* ---- operators -------------------------------------------------------------*/
/* Propagator table (gen_propagator_kernel, device_kernels.cuh:215-234). */
em_status em_propagator(em_plan plan, void* P_dev /* float2[ny][nx] */);
/* Probes in real space at (x[b], y[b]) with defocus c_10 + defocus_shift
 * (Incident_Wave::generate_incident_beam(s), incident_wave.cuh:29-71). */
em_status em_build_probes(em_plan plan, const em_lens* cond, double defocus_shift,
                          const double* x, const double* y, int nprobe, void* psi_dev);
/* Plane wave 1 (set_wave_val_kernel, incident_wave.cuh:45). */
em_status em_plane_wave(em_plan plan, int nprobe, void* psi_dev);
/* Unnormalised 2D FFT of nbatch planes, in place; dir -1 forward, +1 inverse
 * (FFT_Dev::forward/inverse(_batch), fft_dev.cuh:159-190). */
em_status em_fft2d(em_plan plan, void* data_dev, int nbatch, int dir);

/* ---- multislice ------------------------------------------------------------
 * Wave_Function::wave_modulation(_batch) (wave_function.cuh:60-87):
 * psi_dev[B][N] in: incident wave (real space); out: exit wave (Fourier,
 * unnormalised).  Exactly one of V_dev (sigma-scaled FP32 potential
 * [S][N], t = exp(iV)) and t_dev (complex transmission float2[S][N]) is
 * non-NULL.  ckpt_dev (optional, float2[ceil(S/ckpt_every)][B][N]) receives
 * the wave entering slices 0, ckpt_every, 2*ckpt_every, ... (real space), for
 * em_multislice_adjoint; ckpt_every = 1 keeps every slice. */
em_status em_multislice_forward(em_plan plan, const float* V_dev, const void* t_dev, int S,
                                void* psi_dev, int nprobe, void* ckpt_dev, int ckpt_every);
/* Forward multislice where probe b has its own potential stack V_dev[b]
 * (sigma-scaled FP32 [S[b]][N]) and slice count: psi_dev[B][N] in: incident
 * waves (real space); out: exit waves (Fourier, unnormalised), as
 * em_multislice_forward per probe.  All probes share the plan's grid and
 * propagator and run through shared fused passes (a chunk table gives each
 * workgroup its probe's potential slice).  For frozen-phonon configurations,
 * which the reference runs one after another (ED_HRTEM / CBED_CBEI,
 * single_simulator.cuh:60-80, 160-180); no reference counterpart of its own. */
em_status em_multislice_forward_multi(em_plan plan, const float* const* V_dev, const int* S,
                                      void* psi_dev, int nprobe);
/* Adjoint of em_multislice_forward with respect to the transmission (no
 * reference counterpart).  gPsi_dev[B][N] = dL/dRe(Psi) + i dL/dIm(Psi) of
 * the exit waves (overwritten); ckpt_dev and ckpt_every as given to the
 * forward call: waves between checkpoints are recomputed segment by segment
 * into plan-owned workspace.  gt_dev[S][N] receives dL/dRe(t) + i dL/dIm(t)
 * summed over probes; gpsi0_dev (optional, [B][N]) the gradient of the
 * incident waves.  With V_dev, gV_dev (optional, FP32 [S][N]) receives dL/dV
 * of the sigma-scaled potential. */
em_status em_multislice_adjoint(em_plan plan, const float* V_dev, const void* t_dev, int S,
                                const void* ckpt_dev, int ckpt_every, void* gPsi_dev, int nprobe,
                                void* gt_dev, float* gV_dev, void* gpsi0_dev);

/* Integrated detector intensities (integrated_intensity(_batch),
 * wave_function.cuh:135-162; sum_square_over_Det(_batch),
 * device_kernels.cuh:534-579, 625-671): out_dev[d][b] (+)= w * sum |Psi|^2
 * over inner <= |g| <= outer (closed), detector angles in mrad. */
em_status em_detector_sums(em_plan plan, const void* Psi_dev, int nprobe, int ndet,
                           const double* inner_mrad, const double* outer_mrad, float w,
                           int accumulate, float* out_dev);
/* em_detector_sums without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The detector pass's
 * per-tile partials go to a fixed 1 MiB device array (262144 floats) that
 * each GPU holds until the process ends, so nprobe * ndet * (nx / 8) must not
 * exceed 262144 (EM_ERR_INVALID otherwise).  Same results as
 * em_detector_sums.  Calls that overlap on one GPU share the array and race;
 * the plan's device must be current. */
em_status em_detector_sums_noalloc(em_plan plan, const void* Psi_dev, int nprobe, int ndet,
                                   const double* inner_mrad, const double* outer_mrad, float w,
                                   int accumulate, float* out_dev);
/* Coherent objective-lens image (Lens_Modulation::CTF_TEM,
 * lens_modulation.cuh:78-89): I = |IFFT[Psi exp(i chi) A]|^2, or re/im when
 * re_dev/im_dev are given (I_dev NULL); accumulate adds w * value. */
em_status em_ctf_image(em_plan plan, const em_lens* obj, const void* Psi_dev, int nprobe,
                       float* I_dev, float* re_dev, float* im_dev, float w, int accumulate);
/* |Psi|^2 quadrant-shifted (ED / CBED output, single_simulator.cuh:103-110, 189). */
/* Partially coherent image in the quasi-coherent approximation
 * (Lens_Modulation::PCTF_TEM, lens_modulation.cuh:91-127; apply_PCTF_kernel,
 * device_kernels.cuh:504-518): Psi * A(g) * E(g) * e^{i chi(g)}, then IFFT
 * and |.|^2 (or re / im), with the envelope
 *   u = 1 + 2 (pi beta sigma)^2 g^2,
 *   E = exp(-[(pi beta)^2 g^2 (C30 lambda^2 g^2 + C10)^2
 *            + (pi lambda sigma g^2)^2 / 2] / u) / sqrt(u).
 * The reference's envelope writes (C30 lambda^2 g^2 - C10), the opposite
 * defocus sign to its own chi, and takes its defocus half-width for sigma;
 * this form is the Gaussian average of two-beam interference (to first order
 * in the tilt) and agrees with em_tcc_image.  Arguments as em_ctf_image. */
em_status em_pctf_image(em_plan plan, const em_lens* obj, const em_coherence* coh, const void* Psi,
                        int nprobe, float* image, float* re, float* im, float w, int accumulate);
/* Partially coherent image by the transmission cross coefficient
 * (Lens_Modulation::TCC_TEM, lens_modulation.cuh:129-222): the weighted sum
 * of coherent CTF images over Gauss-Hermite nodes, defocus
 * C10 + sqrt(2) sigma t_j (weight w_j / sqrt(pi)) and tilt
 * q0 (t_i, t_k) added to g (weight w_i w_k / pi).  1 to 64 points each. */
em_status em_tcc_image(em_plan plan, const em_lens* obj, const em_coherence* coh, const void* Psi,
                       int nprobe, float* image, float* re, float* im, float w, int accumulate);
em_status em_diffraction(em_plan plan, const void* Psi_dev, int nprobe, float* I_dev, float w,
                         int accumulate);

/* ---- simulations -----------------------------------------------------------
 * Batched STEM over the current specimen (run_STEM_tile_simu's probe-batch
 * loop, scan_simulator.cuh:87-109): potentials of every slice are built once,
 * probes run in batches of max_batch through fused passes captured in a HIP
 * graph, and detector sums land in out_dev[d][ipos] (+)= w * I. */
em_status em_stem(em_plan plan, const em_lens* cond, int npos, const double* x, const double* y,
                  int ndet, const double* inner_mrad, const double* outer_mrad, float w,
                  int accumulate, float* out_dev);
/* Thickness series (run_thickness_simu, thickness_simulator.cuh:11-85):
 * after each slice the wave is band-passed by [g_inner, g_outer] (mrad) and
 * |.|^2 recorded for slices with (k+1) % stride == 0 and the last one.
 * frames_dev[nf][B][N] (+)= w * frame.  psi_dev: incident waves (real space),
 * left holding the real-space wave after the last slice. */
em_status em_thickness(em_plan plan, const float* V_dev, int S, void* psi_dev, int nprobe,
                       double inner_mrad, double outer_mrad, int stride, float w, int accumulate,
                       float* frames_dev, int* n_frames);

/* em_thickness with a potential stack per wave (V_dev[b], sigma-scaled FP32
 * [S][N], one slice count S for all), the row passes sharing one launch over
 * all waves: frozen-phonon configurations of run_thickness_simu
 * (thickness_simulator.cuh:11-85), which the reference runs one after
 * another.  frames_dev[nf][B][N] as em_thickness. */
em_status em_thickness_multi(em_plan plan, const float* const* V_dev, int S, void* psi_dev, int nprobe,
                             double inner_mrad, double outer_mrad, int stride, float w, int accumulate,
                             float* frames_dev, int* n_frames);

/* Ragged batch: requests on different grids run as one HIP graph launch.
 * Requests of one grid size (and band-limit skip range) step through their
 * slices together: each row and propagator column pass of a slice step is one
 * launch over all of their probes, every probe with its own request's
 * potential and propagator; detector passes run per request.  Request r:
 * plan plans[r] (distinct plans), psi_dev[r] incident waves [B_r][N_r]
 * (overwritten), S[r] slices of V_dev[r], detector sums every `stride`
 * slices and after the last one into out_dev[r][nout_r][ndet][B_r] where
 * nout_r = ceil(S_r/stride).  The work is ordered on hip_stream and the call
 * returns without waiting for it.  plans[0] keeps the instantiated graph (up
 * to four request sets): a call with the same buffers, slice counts,
 * detector bands, stride and plan settings re-launches it. */
em_status em_multislice_ragged(int nreq, em_plan* plans, const float* const* V_dev, const int* S,
                               void* const* psi_dev, const int* nprobe, int ndet,
                               const double* inner_mrad, const double* outer_mrad, int stride,
                               float* const* out_dev, void* hip_stream);

/* ---- measurement -----------------------------------------------------------
 * No reference counterpart (the reference has no timing code).  Average
 * device time in ms of the fused row pass (IFFT_x, t = exp(iV), FFT_x) and column
 * pass (FFT_y, P, IFFT_y) over nprobe probes at wave_dev[B][N], each launched
 * iters times between HIP events on the plan's stream.  V_dev: one slice
 * [ny][nx].  The wave is overwritten. */
em_status em_time_passes(em_plan plan, const float* V_dev, void* wave_dev, int nprobe, int iters,
                         float* row_ms, float* col_ms);

#ifdef __cplusplus
}
#endif
#endif /* CUDAEM_AMD_H */
