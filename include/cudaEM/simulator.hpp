// cudaEM/simulator.hpp -- the reference's C++ simulator surface, rebuilt on
// the gfx950 engine (include/cudaem_amd.h).
//
// Names, argument meaning and call order follow the reference so that user
// code written against it compiles and runs unchanged:
//   cudaEM::TEM_simulator<T>        single_simulator.cuh:228-261
//   cudaEM::scan_simulator<T>       scan_simulator.cuh:201-234
//   cudaEM::Thickness_simulator<T>  thickness_simulator.cuh:87-119
//   mulSimulator::execution{TEM,CBED,STEM}   simulator.cu:7-49
// and the host structs the reference uses but does not ship (its
// multislice_params.hpp, atom_coor_data.hpp, output_multislice_data.hpp are
// missing): multislice_params<T>, Atom_Coord_Data<T>, Lens<T>, Detector<T>,
// output_Multi_Data<T>, with the fields the reference reads
// (multislice_params_dev.cuh:88-199, simulator.cu:9-46, lens_dev.cuh:74-131).
//
// Device work runs on one GPU (device 0): GPU_num larger than 1 runs every
// frozen-phonon configuration there (the reference splits them over GPUs,
// multislice_params_dev.cuh:92-97, and averages on the host).
// Errors throw cudaEM::simulation_error (the reference prints and continues).
#pragma once
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cudaEM {

enum eTEM_Sim_Type { eTEMST_HRTEM = 1, eTEMST_ED = 2, eTEMST_CBED = 3, eTEMST_STEM = 4, eTEMST_Thickness = 5 };
enum eIncident_Wave_Type { eIWT_Plane_Wave = 1, eIWT_Convergent_Wave = 2 };
enum ePhonon_Model { ePM_Still_Atom = 1, ePM_Absorptive = 2, ePM_Frozen_Phonon = 3 };
enum eIllumination_Model { eIM_Coherent = 1, eIM_Partial_Coherent = 2, eIM_Trans_Cross_Coef = 3 };
enum eTemporal_Spatial_Incoh { eTSI_Temporal_Spatial = 1, eTSI_Temporal = 2, eTSI_Spatial = 3 };
enum eScanning_Type { eTile_scanning = 1, eGeneral_scanning = 2 };
enum eData_Record { eIntensity = 1, eWave = 2 };
enum eMatch_Border { eMB_Min = 0, eMB_Max = 1, eMB_Middle = 2 };
// output space of ED_HRTEM (single_simulator.cuh:12, 32, 103)
enum eSpace { eS_Real = 1, eS_Reciprocal = 2 };
// 3D potential parameterisations of fine slicing (Vr_dVrir_dev,
// basic_functions_dev.cuh:189-214); values as EM_PT_* of cudaem_amd.h.
enum ePotential_Type {
  ePT_Doyle_0_4 = 1,
  ePT_Peng_0_4 = 2,
  ePT_Peng_0_12 = 3,
  ePT_Kirkland_0_12 = 4,
  ePT_Weickenmeier_0_12 = 5,
  ePT_Lobato_0_12 = 6
};

struct simulation_error : std::runtime_error {
  explicit simulation_error(const std::string& m) : std::runtime_error(m) {}
};

template <class T>
struct Atom_Coord_Data {
  std::vector<T> x, y, z, occ;
  std::vector<int> Z;
  T x_l = 0, y_l = 0;
  std::size_t size() const { return x.size(); }
  void push_back(int Zi, T xi, T yi, T zi, T occi = T(1)) {
    Z.push_back(Zi);
    x.push_back(xi);
    y.push_back(yi);
    z.push_back(zi);
    occ.push_back(occi);
  }
  void assign(const Atom_Coord_Data& o) { *this = o; }
  // translate so that (x0, y0) becomes the origin (scan_simulator.cuh:65, 157)
  void shift_atoms(T x0, T y0) {
    for (T& v : x) v -= x0;
    for (T& v : y) v -= y0;
  }
};

// Keep the atoms inside the closed box [x0, x1] x [y0, y1]; the scanners pass
// their window widened by 2 A on every side (scan_simulator.cuh:61-62, 153-154).
template <class T>
void remove_atoms_outside_xy_range(Atom_Coord_Data<T>& a, T x0, T x1, T y0, T y1) {
  Atom_Coord_Data<T> k;
  k.x_l = a.x_l;
  k.y_l = a.y_l;
  const bool has_occ = a.occ.size() == a.size();
  for (std::size_t i = 0; i < a.size(); ++i)
    if (a.x[i] >= x0 && a.x[i] <= x1 && a.y[i] >= y0 && a.y[i] <= y1) {
      k.Z.push_back(a.Z[i]);
      k.x.push_back(a.x[i]);
      k.y.push_back(a.y[i]);
      k.z.push_back(a.z[i]);
      if (has_occ) k.occ.push_back(a.occ[i]);
    }
  a = std::move(k);
}

// Lens_Dev fields (lens_dev.cuh:12-66): aberrations in Angstrom, angles in
// rad, apertures as semi-angles in mrad.
template <class T>
struct Lens {
  int m = 0;
  T c_10 = 0, c_12 = 0, phi_12 = 0, c_21 = 0, phi_21 = 0, c_23 = 0, phi_23 = 0;
  T c_30 = 0, c_32 = 0, phi_32 = 0, c_34 = 0, phi_34 = 0;
  T c_41 = 0, phi_41 = 0, c_43 = 0, phi_43 = 0, c_45 = 0, phi_45 = 0;
  T c_50 = 0, c_52 = 0, phi_52 = 0, c_54 = 0, phi_54 = 0, c_56 = 0, phi_56 = 0;
  T inner_aper_ang = 0, outer_aper_ang = 0;  // mrad
  // partial coherence (lens_dev.cuh:27-34): standard deviation of the
  // defocus spread (A) and e^-1 half-width semi-angle of the Gaussian source
  // (mrad), with the Gauss-Hermite points of the TCC integrals
  T dsf_sigma = 0;
  int dsf_npoints = 1;
  T ssf_beta = 0;
  int ssf_npoints = 1;
};

// Annular detectors; angles in mrad (wave_function.cuh:135-162 squares the
// converted g_inner / g_outer).
template <class T>
struct Detector {
  std::vector<T> ang_inner, ang_outer;
  std::size_t size() const { return ang_inner.size(); }
};

// 3D potential coefficients of one element (the reference reads them from
// its missing element tables into Vr_coeff, device_containers.cuh:239, 280).
template <class T>
struct Element_Potential {
  int Z = 0;
  T cl[6] = {0, 0, 0, 0, 0, 0};
  T cnl[6] = {1, 1, 1, 1, 1, 1};
};

// A point or direction in 3D (the reference's r3d).
template <class T>
struct r3d {
  T x = 0, y = 0, z = 0;
};

template <class T>
struct multislice_params {
  eTEM_Sim_Type simulation_type = eTEMST_HRTEM;
  int GPU_num = 1;
  bool is_crystal = true;
  // grid (Grid_2d_Dev): pixels, field of view, slice thickness, bwl
  int nx = 0, ny = 0;
  T lx = 0, ly = 0, dz = 0;
  bool bwl = true;
  // beam
  T E_0 = 200;  // kV
  T theta = 0, phi = 0;
  eIncident_Wave_Type iw_type = eIWT_Plane_Wave;
  T incident_wave_x_pos = 0, incident_wave_y_pos = 0;
  eIllumination_Model illumination_model = eIM_Coherent;
  // spreads of partial coherence (host name read at multislice_params_dev.cuh:112)
  eTemporal_Spatial_Incoh temporal_spatial_incoh = eTSI_Temporal_Spatial;
  // specimen
  Atom_Coord_Data<T> atoms;
  // fine slicing (multislice_params_dev.cuh:23-24): each element's 3D
  // potential of type potential_type integrated over every slice it reaches;
  // elements without an entry in element_potentials use Moliere's screening
  // in the Kirkland form
  bool enable_fine_slicing = false;
  ePotential_Type potential_type = ePT_Kirkland_0_12;
  std::vector<Element_Potential<T>> element_potentials;
  int fine_slicing_nq = 128;  // Gauss-Legendre points per slice
  // phonons: Gaussian displacements of rms pn_sigma along enabled pn_dim
  ePhonon_Model phonon_model = ePM_Still_Atom;
  int pn_nconf = 1;
  unsigned long long pn_seed = 1;
  T pn_sigma = 0;
  bool pn_dim[3] = {true, true, false};
  // lenses and detectors
  Lens<T> obj_lens, cond_lens;
  Detector<T> detector;
  eData_Record data_record = eIntensity;
  // STEM scan: STEM_col x STEM_row positions over the normalised range
  // [x0, y0, x1, y1] of the field (set_Scanning, simulator.cu:41-46)
  eScanning_Type stem_scanning = eTile_scanning;
  int STEM_col = 0, STEM_row = 0;
  T scan_range[4] = {0, 0, 1, 1};
  int probe_batch = 256;  // probes per batched launch (c_Probe_Batch)
  // thickness series: frames every thickness_stride slices through
  // detector thickness_det (the reference uses detector 1)
  int thickness_stride = 1;
  int thickness_det = 1;
  // specimen rotation (multislice_params_dev.cuh:35-37, 119-121): carried
  // like the reference's shipped code carries them; none of its kernels
  // applies a rotation, and neither does this build
  int nrot = 1;
  r3d<T> spec_rot_u0{0, 0, 1};
  r3d<T> spec_rot_center_p{0, 0, 0};

  void assign(const multislice_params& o) { *this = o; }
  void validate_parameters() const;
  void initial_OL() { iw_type = eIWT_Plane_Wave; }
  void initial_CL() { iw_type = eIWT_Convergent_Wave; }
  void set_Scanning(const std::vector<float>& r) {
    if (r.size() != 4) throw simulation_error("scan range needs 4 values");
    for (int i = 0; i < 4; ++i) scan_range[i] = (T)r[i];
  }
  bool frozen_phonon_enable() const { return phonon_model == ePM_Frozen_Phonon; }
  bool scanning_enable() const { return simulation_type == eTEMST_STEM; }
  // batched probe launches (tile scanning with more than one probe a batch)
  bool batch_enable() const { return stem_scanning == eTile_scanning && probe_batch > 1; }
  int output_scan_pos_num() const { return STEM_col * STEM_row; }
  // accessors the reference's device parameters read (multislice_params_dev.cuh:88-199)
  int output_STEM_Col() const { return STEM_col; }
  int output_STEM_Row() const { return STEM_row; }
  unsigned long long output_phonn_seed() const { return pn_seed; }
  eIncident_Wave_Type output_incident_wave_type() const { return iw_type; }
  const Atom_Coord_Data<T>& output_Atom_data() const { return atoms; }
  T output_scan_interval_x() const { return STEM_col > 0 ? (scan_range[2] - scan_range[0]) * lx / STEM_col : T(0); }
  T output_scan_interval_y() const { return STEM_row > 0 ? (scan_range[3] - scan_range[1]) * ly / STEM_row : T(0); }
  // relativistic wavelength (A), gamma and interaction constant (rad/(V A)) at E_0
  T output_lamda() const;
  T output_gama() const;
  T output_sigma() const;
};

template <class T>
struct output_Multi_Data {
  int row = 0, col = 0;
  T interval_x = 0, interval_y = 0;
  std::vector<T> image_intensity, image_real, image_imag;
  std::vector<T> STEM_intensity;     // [iDet][y][x] (tile) or [iDet][ipos] (general)
  std::vector<T> thickness_frames;   // [frame][pixel]
  T total_intensity = 0;             // incident intensity, sum over g of |Psi_0(g)|^2
  void init(int row_i, int col_i, int ndet, bool full_wave) {
    row = row_i;
    col = col_i;
    if (ndet > 0) {
      STEM_intensity.assign((std::size_t)row * col * ndet, T(0));
    } else if (full_wave) {
      image_real.assign((std::size_t)row * col, T(0));
      image_imag.assign((std::size_t)row * col, T(0));
    } else {
      image_intensity.assign((std::size_t)row * col, T(0));
    }
  }
};

template <class T>
class TEM_simulator {
 public:
  void set_input_data(multislice_params<T>* p, output_Multi_Data<T>* o) {
    multisliceparam_host_ptr = p;
    output_mulData_ptr = o;
  }
  void TEM_multiGPU();
  output_Multi_Data<T>* output_mulData_ptr = nullptr;

 private:
  multislice_params<T>* multisliceparam_host_ptr = nullptr;
};

template <class T>
class scan_simulator {
 public:
  void set_input_data(multislice_params<T>* p, output_Multi_Data<T>* o) {
    multisliceparam_host_ptr = p;
    output_mulData_ptr = o;
  }
  void STEM_multiGPU();
  output_Multi_Data<T>* output_mulData_ptr = nullptr;

 private:
  multislice_params<T>* multisliceparam_host_ptr = nullptr;
};

template <class T>
class Thickness_simulator {
 public:
  void set_input_data(multislice_params<T>* p, output_Multi_Data<T>* o) {
    multisliceparam_host_ptr = p;
    output_mulData_ptr = o;
  }
  void Thickness_multiGPU();
  output_Multi_Data<T>* output_mulData_ptr = nullptr;

 private:
  multislice_params<T>* multisliceparam_host_ptr = nullptr;
};

// Per-device entry points of the reference (single_simulator.cuh:207,
// scan_simulator.cuh:184, thickness_simulator.cuh:11).
template <class T>
void run_simu(multislice_params<T>* p, int GPU_ID, output_Multi_Data<T>* o);
template <class T>
void run_STEM_simu(multislice_params<T>* p, int GPU_ID, output_Multi_Data<T>* o);
template <class T>
void run_thickness_simu(multislice_params<T>* p, int GPU_ID, output_Multi_Data<T>* o);

// Seeded frozen-phonon displacement (the reference's missing
// vibrate_atom_positions): splitmix64 uniforms, Box-Muller normals consumed
// x, y, z per atom, configuration iconf.
template <class T>
void vibrate_atom_positions(const Atom_Coord_Data<T>& in, Atom_Coord_Data<T>& out, unsigned long long seed,
                            int iconf, T sigma, const bool dims[3]);

// Raw result files (the reference's rawdata_io.hpp is missing): the values
// of v in native byte order, no header.
template <class T>
void write_raw_data(const std::string& path, const std::vector<T>& v) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw simulation_error("cannot write " + path);
  const size_t n = v.empty() ? 0 : std::fwrite(v.data(), sizeof(T), v.size(), f);
  const bool closed = std::fclose(f) == 0;
  if (n != v.size() || !closed) throw simulation_error("short write to " + path);
}

template <class T>
std::vector<T> read_raw_data(const std::string& path) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw simulation_error("cannot read " + path);
  std::vector<T> v;
  T buf[4096];
  size_t n;
  while ((n = std::fread(buf, sizeof(T), 4096, f)) > 0) v.insert(v.end(), buf, buf + n);
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  if (bad) throw simulation_error("read error in " + path);
  return v;
}

// Atom list as text (the reference's iolib.hpp is missing): one atom per
// line, "Z x y z [occ]" with positions in Angstrom; blank lines and text
// after '#' are ignored.  x_l / y_l are left for the caller to set.
template <class T>
Atom_Coord_Data<T> read_atoms_text(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw simulation_error("cannot read " + path);
  Atom_Coord_Data<T> at;
  std::string line;
  int lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.erase(hash);
    std::istringstream ss(line);
    int Z;
    double x, y, z, occ = 1.0;
    if (!(ss >> Z)) continue;  // blank line
    if (!(ss >> x >> y >> z)) throw simulation_error(path + ":" + std::to_string(lineno) + ": expected Z x y z");
    if (!(ss >> occ)) occ = 1.0;
    if (Z <= 0 || Z > 103) throw simulation_error(path + ":" + std::to_string(lineno) + ": atomic number");
    at.push_back(Z, (T)x, (T)y, (T)z, (T)occ);
  }
  return at;
}

}  // namespace cudaEM

class mulSimulator {
 public:
  void executionTEM(const cudaEM::multislice_params<float>& in, cudaEM::output_Multi_Data<float>* out);
  void executionCBED(const cudaEM::multislice_params<float>& in, cudaEM::output_Multi_Data<float>* out);
  void executionSTEM(const cudaEM::multislice_params<float>& in, cudaEM::output_Multi_Data<float>* out);
  void executionThickness(const cudaEM::multislice_params<float>& in, cudaEM::output_Multi_Data<float>* out);
};
