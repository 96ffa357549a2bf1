// nbody_io.hpp -- the reference's file formats (inputs and outputs), kept
// byte-compatible so this driver is a drop-in for nbody/src/nbody.cu.
#pragma once

#include <array>
#include <cstdio>
#include <string>
#include <vector>

namespace khrr_host {

// parse.cuh:19-62
struct GalaxyProperties {
  int n_star_particles = 0;
  int n_dark_particles = 0;
  double mass_stars = 0.;
  double mass_dark = 0.;
  double pos_center[3] = {0., 0., 0.};
  double vel_center[3] = {0., 0., 0.};
  double angle = 0.;
  double eps_stars = 0.;
  double eps_dark = 0.;

  int get_n_star_particles() const { return mass_stars > 0. ? n_star_particles : 0; }
  int get_n_dark_particles() const { return mass_dark > 0. ? n_dark_particles : 0; }
  double get_mass_star_particle() const { return mass_stars / get_n_star_particles(); }
  double get_mass_dark_particle() const { return mass_dark / get_n_dark_particles(); }
  double get_angle_rad() const;
};

struct StartInfo {  // __start_nbody.ini, nbody.cu:961-984
  int i_cont = 0;
  double tmax = 0.;
  double dtsave = 0.;
};

struct GravParams {  // __gr_par.ini, nbody.cu:1080-1106
  double Mh = 0, a = 0, Rh = 0, Mb = 0, b = 0, Rb = 0, eps = 0, dtgrav = 0.001, K_m = 0, K_r = 0;
};

// Host particle state in the reference's order (all stars of all galaxies,
// then all DM), as vec3 planes.
struct Particles {
  int n = 0;
  int n_star = 0;
  std::vector<double> pos, vel;  // 3*n planes
  std::vector<double> mass;      // n
  std::vector<double> eps2;      // n (eps^2, as the reference's eps2_host)
};

// Each reader returns false and fills `err` on failure (the reference
// prints and calls exit(0); this driver reports and returns).
bool read_galaxy_properties(const std::string& path, std::vector<GalaxyProperties>& out,
                            std::string& err);
bool read_start_info(const std::string& path, StartInfo& out, std::string& err);
bool read_gr_par(const std::string& path, GravParams& out, std::string& err);
// nbody.cu:1315-1421: start_S{k}.txt / start_DM{k}.txt with the per-galaxy
// rotation and translation applied in place (z uses the rotated x).
bool load_start_particles(const std::string& ini_dir, const std::vector<GalaxyProperties>& g,
                          Particles& p, std::string& err);

// nbody.cu:726-755: {name}_{it:5}.bin = int count, double t, then
// x y z vx vy vz doubles per particle.
bool write_particles_bin(const std::string& bin_dir, const char* name, int i0, int icount,
                         const Particles& p, int it, double t);
// "{:5}" of std::format for an int: right-aligned in width 5.
std::string width5(int v);

// Conservation-law series of result() (nbody.cu:897-957) and LIE_0.bin
// (:844-848, written as L0, P0, E0 and 16 zero bytes).
struct Diag {
  double Ek, Ep, P[3], L[3];
  double Vr_max, Vfi_max, Vz_max, R_max, Z_max;
};
class ResultWriter {
 public:
  ResultWriter(std::string bin_dir, std::string out_dir)
      : bin_dir_(std::move(bin_dir)), out_dir_(std::move(out_dir)) {}
  // it == 0 records the initial values and truncates the CSV files.
  bool write(const Diag& d, int it, double t, int it_all);
  void set_initial(const double L0[3], const double P0[3], double E0);
  double E0() const { return E0_; }

 private:
  std::string bin_dir_, out_dir_;
  double L0_[3] = {0, 0, 0}, P0_[3] = {0, 0, 0}, E0_ = 0;
};

// out/sample.csv row for particle 1 (nbody.cu:1669-1699).
bool write_sample(const std::string& out_dir, int it, int it_all, double t, const double pos[3],
                  const double vel[3], const double acc[3], double phi);

// "x, <value>" rows along the x line of cells at iy = iz = NX/2, x = -sim_l +
// i*dx in std::format's shortest "{}" form and the value with `decimals`
// fixed decimals (print_cell_phi nbody.cu:704-724, the slices of
// convert_particles :428-645 and check_mass :107-112).
bool write_x_slice(const std::string& path, const char* header, int nx, double sim_l, double dx,
                   const std::vector<double>& cell_value, int decimals);
// out/phi/phi_{:5}.csv: x-slice of the cell potential (print_cell_phi).
bool write_cell_phi(const std::string& out_dir, int iter, int nx, double sim_l, double dx,
                    const std::vector<double>& cell_phi);

// Debug dumps of convert_particles (nbody.cu:287-701) into data/nbody/debug:
// "i, acc" rows of one value per body, seven decimals (:622-697).
bool write_index_list(const std::string& path, const std::vector<double>& values);
// particle_counts.txt: "xi, counts" along the x line (:412-432).
bool write_count_slice(const std::string& path, int nx, const std::vector<int>& count);
// particles_in_cell.txt (count, cells with that count, largest count first)
// and particles_in_cell_partial_sum.txt (running total of bodies over those
// rows from the second on) of check_particles_distribution (:68-96).
bool write_count_distribution(const std::string& dir, const std::vector<int>& count);

// %g-style text of a double as std::ostream's default float format prints it.
std::string ostream_double(double v);

}  // namespace khrr_host
