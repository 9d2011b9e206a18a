// Drives the C++ simulator surface (include/cudaEM/simulator.hpp) the way
// user code of the reference does, and writes the results as raw float32
// files for tests/test_gpu_simulators.py to compare with the oracle.
//
//   test_simulators vibrate <out_dir>   phonon displacements only (no GPU)
//   test_simulators io <out_dir>        atoms.txt -> atoms.f32 through the I/O helpers (no GPU)
//   test_simulators params <out_dir>    parameter accessors -> params.f64 (no GPU)
//   test_simulators cut <out_dir>       atom window cut and shift -> cut.f64 (no GPU)
//   test_simulators run <out_dir>       HRTEM, ED, CBED, STEM (tile and
//                                       general scanning), thickness series
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cudaEM/simulator.hpp"

using namespace cudaEM;

static void write(const std::string& path, const std::vector<float>& v) { write_raw_data(path, v); }

// SrTiO3 perovskite, n x n x nz cells (a = 3.905 A)
static Atom_Coord_Data<float> srtio3(int n, int nz) {
  const double a = 3.905;
  const double basis[5][4] = {{38, 0, 0, 0}, {22, .5, .5, .5}, {8, .5, .5, 0}, {8, .5, 0, .5}, {8, 0, .5, .5}};
  Atom_Coord_Data<float> at;
  for (int b = 0; b < 5; ++b)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        for (int k = 0; k < nz; ++k)
          at.push_back((int)basis[b][0], (float)((i + basis[b][1]) * a), (float)((j + basis[b][2]) * a),
                       (float)((k + basis[b][3]) * a));
  at.x_l = at.y_l = (float)(n * a);
  return at;
}

static multislice_params<float> base(int n, int cells, int nz) {
  multislice_params<float> p;
  p.atoms = srtio3(cells, nz);
  p.nx = p.ny = n;
  p.lx = p.ly = cells * 3.905f;
  p.dz = 3.905f / 2;
  p.E_0 = 300;
  p.cond_lens.outer_aper_ang = 20;
  p.detector.ang_inner = {0, 40};
  p.detector.ang_outer = {20, 120};
  return p;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s vibrate|run <out_dir>\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1], out = argv[2];
  try {
    if (mode == "config2") {
      // Plain STEM ADF as the reference's general scanner runs it: 512^2,
      // SrTiO3 10 x 10 x 40 cells, 80 slices, 300 kV, 25 mrad, Cs = 0,
      // defocus 0, detector 60-200 mrad; positions (5, 3), (16, 3), (5, 16),
      // (16, 16) of the 32 x 32 raster over the field.  mulSimulator always
      // scans the whole field (simulator.cu:40-46), so a subset of the raster
      // goes through scan_simulator with its own scan range, as in
      // executionSTEM otherwise.
      multislice_params<float> p = base(512, 10, 40);
      p.simulation_type = eTEMST_STEM;
      p.cond_lens.outer_aper_ang = 25;
      p.detector.ang_inner = {60};
      p.detector.ang_outer = {200};
      p.STEM_col = 2;
      p.STEM_row = 2;
      p.stem_scanning = eGeneral_scanning;
      p.is_crystal = false;
      p.validate_parameters();
      p.initial_CL();
      p.set_Scanning({5.f / 32, 3.f / 32, 27.f / 32, 29.f / 32});
      output_Multi_Data<float> o;
      scan_simulator<float> sim;
      sim.set_input_data(&p, &o);
      sim.STEM_multiGPU();
      write(out + "/config2_stem.f32", o.STEM_intensity);
      write(out + "/config2_total.f32", {o.total_intensity});
      std::printf("config2 %zu total %g\n", o.STEM_intensity.size(), o.total_intensity);
      return 0;
    }
    if (mode == "io") {  // text atoms in, raw float32 out and back (no GPU)
      const Atom_Coord_Data<float> at = read_atoms_text<float>(out + "/atoms.txt");
      std::vector<float> r;
      for (size_t i = 0; i < at.size(); ++i) {
        r.push_back((float)at.Z[i]);
        r.push_back(at.x[i]);
        r.push_back(at.y[i]);
        r.push_back(at.z[i]);
        r.push_back(at.occ[i]);
      }
      write(out + "/atoms.f32", r);
      if (read_raw_data<float>(out + "/atoms.f32") != r) throw simulation_error("raw round trip");
      return 0;
    }
    if (mode == "cut") {  // remove_atoms_outside_xy_range + shift_atoms (no GPU)
      Atom_Coord_Data<double> at;
      for (int i = 0; i < 40; ++i) at.push_back(6 + i % 3, -3.0 + 0.5 * i, 7.0 - 0.35 * i, 0.1 * i, 0.5 + 0.01 * i);
      remove_atoms_outside_xy_range(at, -1.0 - 2.0, 9.0 + 2.0, -2.0 - 2.0, 4.0 + 2.0);
      at.shift_atoms(-1.0, -2.0);
      std::vector<double> r;
      for (size_t i = 0; i < at.size(); ++i) {
        r.push_back(at.Z[i]);
        r.push_back(at.x[i]);
        r.push_back(at.y[i]);
        r.push_back(at.z[i]);
        r.push_back(at.occ[i]);
      }
      write_raw_data(out + "/cut.f64", r);
      return 0;
    }
    if (mode == "params") {  // the reference's parameter names and accessors (no GPU)
      multislice_params<double> p;
      p.E_0 = 300;
      p.temporal_spatial_incoh = eTSI_Temporal;
      p.nrot = 1;
      p.spec_rot_u0 = r3d<double>{0, 0, 1};
      p.spec_rot_center_p = r3d<double>{1, 2, 3};
      p.lx = 40;
      p.ly = 30;
      p.STEM_col = 8;
      p.STEM_row = 6;
      p.set_Scanning({0.f, 0.f, 0.5f, 1.f});
      p.pn_seed = 42;
      p.initial_CL();
      const std::vector<double> r = {p.output_lamda(), p.output_gama(), p.output_sigma(),
                                     (double)p.output_STEM_Col(), (double)p.output_STEM_Row(),
                                     p.output_scan_interval_x(), p.output_scan_interval_y(),
                                     (double)p.output_phonn_seed(),
                                     p.output_incident_wave_type() == eIWT_Convergent_Wave ? 1.0 : 0.0,
                                     p.batch_enable() ? 1.0 : 0.0, (double)p.output_scan_pos_num()};
      write_raw_data(out + "/params.f64", r);
      return 0;
    }
    if (mode == "vibrate") {
      Atom_Coord_Data<double> at, v;
      for (int i = 0; i < 7; ++i) at.push_back(8, 0.5 * i, 1.0 * i, 0.25 * i);
      const bool dims[3] = {true, true, true};
      vibrate_atom_positions(at, v, 1234ull, 3, 0.08, dims);
      std::vector<float> r;
      for (size_t i = 0; i < v.size(); ++i) {
        r.push_back((float)(v.x[i] - at.x[i]));
        r.push_back((float)(v.y[i] - at.y[i]));
        r.push_back((float)(v.z[i] - at.z[i]));
      }
      write(out + "/vibrate.f32", r);
      return 0;
    }
    mulSimulator sim;
    {  // HRTEM image (plane wave, Cs = 1 mm, defocus -50 nm)
      multislice_params<float> p = base(128, 3, 4);
      p.simulation_type = eTEMST_HRTEM;
      p.obj_lens.c_30 = 1e7f;
      p.obj_lens.c_10 = -500.f;
      output_Multi_Data<float> o;
      sim.executionTEM(p, &o);
      write(out + "/hrtem.f32", o.image_intensity);
      std::printf("hrtem %zu\n", o.image_intensity.size());
    }
    for (int model = 2; model <= 3; ++model) {  // HRTEM, partial coherence: PCTF, TCC
      multislice_params<float> p = base(128, 3, 4);
      p.simulation_type = eTEMST_HRTEM;
      p.illumination_model = (eIllumination_Model)model;
      p.obj_lens.c_30 = 1e7f;
      p.obj_lens.c_10 = -500.f;
      p.obj_lens.outer_aper_ang = 30.f;
      p.obj_lens.dsf_sigma = 40.f;
      p.obj_lens.dsf_npoints = 5;
      p.obj_lens.ssf_beta = 0.3f;
      p.obj_lens.ssf_npoints = 3;
      output_Multi_Data<float> o;
      sim.executionTEM(p, &o);
      write(out + (model == 2 ? "/hrtem_pctf.f32" : "/hrtem_tcc.f32"), o.image_intensity);
      std::printf("hrtem partial coherence %d %zu\n", model, o.image_intensity.size());
    }
    {  // ED, full wave
      multislice_params<float> p = base(128, 3, 4);
      p.simulation_type = eTEMST_ED;
      p.data_record = eWave;
      output_Multi_Data<float> o;
      sim.executionTEM(p, &o);
      write(out + "/ed_re.f32", o.image_real);
      write(out + "/ed_im.f32", o.image_imag);
      std::printf("ed %zu\n", o.image_real.size());
    }
    {  // ED, full wave, fine slicing; Sr in the Lobato form, O and Ti default
      multislice_params<float> p = base(128, 3, 4);
      p.simulation_type = eTEMST_ED;
      p.data_record = eWave;
      p.enable_fine_slicing = true;
      p.fine_slicing_nq = 64;
      p.potential_type = ePT_Lobato_0_12;
      Element_Potential<float> sr;
      sr.Z = 38;
      const float cl[5] = {1.f, 3.f, 8.f, 15.f, 4.f}, cnl[5] = {4.f, 6.f, 10.f, 20.f, 40.f};
      for (int i = 0; i < 5; ++i) {
        sr.cl[i] = cl[i];
        sr.cnl[i] = cnl[i];
      }
      p.element_potentials.push_back(sr);
      output_Multi_Data<float> o;
      sim.executionTEM(p, &o);
      write(out + "/ed_fine_re.f32", o.image_real);
      write(out + "/ed_fine_im.f32", o.image_imag);
      std::printf("ed_fine %zu\n", o.image_real.size());
    }
    {  // CBED at (5, 6), two frozen-phonon configurations
      multislice_params<float> p = base(128, 3, 4);
      p.simulation_type = eTEMST_CBED;
      p.incident_wave_x_pos = 5.f;
      p.incident_wave_y_pos = 6.f;
      p.phonon_model = ePM_Frozen_Phonon;
      p.pn_nconf = 2;
      p.pn_seed = 77;
      p.pn_sigma = 0.08f;
      output_Multi_Data<float> o;
      sim.executionCBED(p, &o);
      write(out + "/cbed.f32", o.image_intensity);
      std::printf("cbed %zu\n", o.image_intensity.size());
    }
    for (int general = 0; general < 2; ++general) {  // STEM, 4 x 3 scan
      multislice_params<float> p = base(128, 3, 4);
      p.simulation_type = eTEMST_STEM;
      p.STEM_col = 4;
      p.STEM_row = 3;
      p.probe_batch = 5;
      p.stem_scanning = general ? eGeneral_scanning : eTile_scanning;
      output_Multi_Data<float> o;
      sim.executionSTEM(p, &o);
      write(out + (general ? "/stem_general.f32" : "/stem_tile.f32"), o.STEM_intensity);
      write(out + (general ? "/stem_general_total.f32" : "/stem_tile_total.f32"), {o.total_intensity});
      std::printf("stem %d %zu total %g\n", general, o.STEM_intensity.size(), o.total_intensity);
    }
    {  // thickness series through detector 1, every 2 slices
      multislice_params<float> p = base(128, 3, 4);
      p.simulation_type = eTEMST_Thickness;
      p.iw_type = eIWT_Convergent_Wave;
      p.incident_wave_x_pos = 2.f;
      p.incident_wave_y_pos = 3.f;
      p.thickness_stride = 2;
      output_Multi_Data<float> o;
      sim.executionThickness(p, &o);
      write(out + "/thickness.f32", o.thickness_frames);
      std::printf("thickness %zu\n", o.thickness_frames.size());
    }
    {  // thickness series with three frozen-phonon configurations (one batch)
      multislice_params<float> p = base(128, 3, 4);
      p.simulation_type = eTEMST_Thickness;
      p.iw_type = eIWT_Convergent_Wave;
      p.incident_wave_x_pos = 2.f;
      p.incident_wave_y_pos = 3.f;
      p.thickness_stride = 3;
      p.phonon_model = ePM_Frozen_Phonon;
      p.pn_nconf = 3;
      p.pn_seed = 91;
      p.pn_sigma = 0.08f;
      output_Multi_Data<float> o;
      sim.executionThickness(p, &o);
      write(out + "/thickness_fp.f32", o.thickness_frames);
      std::printf("thickness_fp %zu\n", o.thickness_frames.size());
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
