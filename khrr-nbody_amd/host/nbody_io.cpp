// nbody_io.cpp -- readers and writers of the reference's input and output
// files (see nbody_io.hpp for the format citations).
#include "nbody_io.hpp"

#include <charconv>
#include <cmath>
#include <cstring>
#include <map>
#include <sys/stat.h>

namespace khrr_host {

namespace {
constexpr double kPi = 3.14159265358979;  // common.cuh:7
constexpr double kDeg2Rad = kPi / 180.0;  // common.cuh:10

struct File {
  FILE* f = nullptr;
  File() = default;
  File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {}
  File(const File&) = delete;
  File& operator=(const File&) = delete;
  ~File() {
    if (f) std::fclose(f);
  }
};

std::string join(const std::string& dir, const std::string& name) {
  if (dir.empty()) return name;
  return dir.back() == '/' ? dir + name : dir + "/" + name;
}
}  // namespace

double GalaxyProperties::get_angle_rad() const { return kDeg2Rad * angle; }

std::string width5(int v) {
  char buf[32];
  std::snprintf(buf, sizeof(buf), "%5d", v);
  return buf;
}

std::string ostream_double(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%g", v);
  return buf;
}

// parse.cuh:320-367 (same fscanf patterns)
bool read_galaxy_properties(const std::string& path, std::vector<GalaxyProperties>& out,
                            std::string& err) {
  File fp(path, "r");
  if (!fp.f) {
    err = "Can't open galaxy properties file: \"" + path + "\"";
    return false;
  }
  char temp[256];
  int count = 0, id = 0;
  if (std::fscanf(fp.f, "%d %255[^\n]", &count, temp) < 1 || count < 0) {
    err = "bad galaxy count in " + path;
    return false;
  }
  out.assign(count, GalaxyProperties{});
  for (auto& g : out) {
    int ok = 0;
    ok += std::fscanf(fp.f, "%d %255[^\n]", &id, temp) >= 1;
    ok += std::fscanf(fp.f, "%d,%d %255[^\n]", &g.n_star_particles, &g.n_dark_particles, temp) >= 2;
    ok += std::fscanf(fp.f, "%lf,%lf %255[^\n]", &g.mass_stars, &g.mass_dark, temp) >= 2;
    ok += std::fscanf(fp.f, "%lf,%lf %255[^\n]", &g.eps_stars, &g.eps_dark, temp) >= 2;
    ok += std::fscanf(fp.f, "%lf %255[^\n]", &g.angle, temp) >= 1;
    ok += std::fscanf(fp.f, "%lf,%lf,%lf %255[^\n]", &g.pos_center[0], &g.pos_center[1],
                      &g.pos_center[2], temp) >= 3;
    ok += std::fscanf(fp.f, "%lf,%lf,%lf %255[^\n]", &g.vel_center[0], &g.vel_center[1],
                      &g.vel_center[2], temp) >= 3;
    if (ok != 7) {
      err = "truncated galaxy block in " + path;
      return false;
    }
  }
  return true;
}

// nbody.cu:961-984
bool read_start_info(const std::string& path, StartInfo& out, std::string& err) {
  File fp(path, "r");
  if (!fp.f) {
    err = "can't open " + path;
    return false;
  }
  char temp[512];
  int ok = std::fscanf(fp.f, "%d  %511[^\n]", &out.i_cont, temp) >= 1;
  ok += std::fscanf(fp.f, "%lf  %511[^\n]", &out.tmax, temp) >= 1;
  ok += std::fscanf(fp.f, "%lf  %511[^\n]", &out.dtsave, temp) >= 1;
  if (ok != 3) {
    err = "truncated " + path;
    return false;
  }
  return true;
}

// nbody.cu:1080-1106 (eps is kept as read; the caller squares it)
bool read_gr_par(const std::string& path, GravParams& g, std::string& err) {
  File fp(path, "r");
  if (!fp.f) {
    err = "can't open " + path;
    return false;
  }
  char temp[512];
  double* dst[] = {&g.Mh, &g.a, &g.Rh, &g.Mb, &g.b, &g.Rb, &g.eps, &g.dtgrav, &g.K_m, &g.K_r};
  for (double* d : dst) {
    if (std::fscanf(fp.f, "%lf  %511[^\n]", d, temp) < 1) {
      err = "truncated " + path;
      return false;
    }
  }
  return true;
}

namespace {
// One component file: header "N t", then N lines of x y z vx vy vz; the
// galaxy's rotation about y and translation, applied in place as the
// reference does (the z update reads the already-rotated x, nbody.cu:1344-1358).
bool load_component(const std::string& path, int count, const GalaxyProperties& g, double mass,
                    double eps, int n0, Particles& p, std::string& err) {
  File fp(path, "r");
  if (!fp.f) {
    err = "Error OF -- " + path;
    return false;
  }
  int itmp = 0;
  double rtmp = 0.0;
  if (std::fscanf(fp.f, "%d %lf", &itmp, &rtmp) != 2) {
    err = "bad header in " + path;
    return false;
  }
  const size_t N = (size_t)p.n;
  const double ca = std::cos(g.get_angle_rad()), sa = std::sin(g.get_angle_rad());
  for (int i = n0; i < n0 + count; ++i) {
    double x, y, z, vx, vy, vz;
    if (std::fscanf(fp.f, "%lf %lf %lf %lf %lf %lf", &x, &y, &z, &vx, &vy, &vz) != 6) {
      err = "truncated " + path;
      return false;
    }
    p.mass[i] = mass;
    x = g.pos_center[0] + x * ca + z * sa;
    y += g.pos_center[1];
    z = g.pos_center[2] + z * ca - x * sa;
    vx = g.vel_center[0] + vx * ca + vz * sa;
    vy += g.vel_center[1];
    vz = g.vel_center[2] + vz * ca - vx * sa;
    p.pos[i] = x;
    p.pos[N + i] = y;
    p.pos[2 * N + i] = z;
    p.vel[i] = vx;
    p.vel[N + i] = vy;
    p.vel[2 * N + i] = vz;
    p.eps2[i] = eps * eps;
  }
  return true;
}
}  // namespace

bool load_start_particles(const std::string& ini_dir, const std::vector<GalaxyProperties>& gal,
                          Particles& p, std::string& err) {
  int ns = 0, nd = 0;
  for (auto& g : gal) {
    ns += g.n_star_particles;  // totals as nbody.cu:1041-1057
    nd += g.n_dark_particles;
  }
  p.n = ns + nd;
  p.n_star = ns;
  p.pos.assign(3 * (size_t)p.n, 0.0);
  p.vel.assign(3 * (size_t)p.n, 0.0);
  p.mass.assign(p.n, 0.0);
  p.eps2.assign(p.n, 0.0);
  int n0 = 0;
  for (size_t k = 0; k < gal.size(); ++k) {
    const auto& g = gal[k];
    if (g.get_n_star_particles() > 0) {
      if (!load_component(join(ini_dir, "start_S" + std::to_string(k) + ".txt"),
                          g.get_n_star_particles(), g, g.get_mass_star_particle(), g.eps_stars,
                          n0, p, err))
        return false;
      n0 += g.get_n_star_particles();
    }
  }
  for (size_t k = 0; k < gal.size(); ++k) {
    const auto& g = gal[k];
    if (g.get_n_dark_particles() > 0) {
      if (!load_component(join(ini_dir, "start_DM" + std::to_string(k) + ".txt"),
                          g.get_n_dark_particles(), g, g.get_mass_dark_particle(), g.eps_dark,
                          n0, p, err))
        return false;
      n0 += g.get_n_dark_particles();
    }
  }
  return true;
}

bool write_particles_bin(const std::string& bin_dir, const char* name, int i0, int icount,
                         const Particles& p, int it, double t) {
  const std::string path = join(bin_dir, std::string(name) + "_" + width5(it) + ".bin");
  File fp(path, "wb");
  if (!fp.f) return false;
  std::fwrite(&icount, sizeof(int), 1, fp.f);
  std::fwrite(&t, sizeof(double), 1, fp.f);
  const size_t N = (size_t)p.n;
  std::vector<double> rec(6 * (size_t)icount);
  for (int i = i0; i < i0 + icount; ++i) {
    double* r = &rec[6 * (size_t)(i - i0)];
    r[0] = p.pos[i];
    r[1] = p.pos[N + i];
    r[2] = p.pos[2 * N + i];
    r[3] = p.vel[i];
    r[4] = p.vel[N + i];
    r[5] = p.vel[2 * N + i];
  }
  return std::fwrite(rec.data(), sizeof(double), rec.size(), fp.f) == rec.size();
}

void ResultWriter::set_initial(const double L0[3], const double P0[3], double E0) {
  for (int k = 0; k < 3; ++k) {
    L0_[k] = L0[k];
    P0_[k] = P0[k];
  }
  E0_ = E0;
}

bool ResultWriter::write(const Diag& d, int it, double t, int it_all) {
  const double E = 0.5 * (d.Ek + d.Ep);
  if (it == 0) {
    set_initial(d.L, d.P, E);
    File fp(join(bin_dir_, "LIE_0.bin"), "wb");
    if (!fp.f) return false;
    const double tail[2] = {0.0, 0.0};
    std::fwrite(L0_, sizeof(double), 3, fp.f);
    std::fwrite(P0_, sizeof(double), 3, fp.f);
    std::fwrite(&E0_, sizeof(double), 1, fp.f);
    std::fwrite(tail, sizeof(double), 2, fp.f);
  }
  const double LL = std::sqrt(d.L[0] * d.L[0] + d.L[1] * d.L[1] + d.L[2] * d.L[2]);
  const double LL0 = std::sqrt(L0_[0] * L0_[0] + L0_[1] * L0_[1] + L0_[2] * L0_[2]);
  const double Imp = std::sqrt(d.P[0] * d.P[0] + d.P[1] * d.P[1] + d.P[2] * d.P[2]);
  const double Imp0 = std::sqrt(P0_[0] * P0_[0] + P0_[1] * P0_[1] + P0_[2] * P0_[2]);
  std::printf("           ***Star***\n R_max = %g  Z_max = %g \n", d.R_max, d.Z_max);
  std::printf("Vr_max = %g  Vfi_max = %g  Vz_max = %g \n", d.Vr_max, d.Vfi_max, d.Vz_max);
  std::printf("           ***Conservation Laws***\n");
  std::printf("----Imp0--- = %g; %g; %g  \n", P0_[0], P0_[1], P0_[2]);
  std::printf("----dImp--- = %g; %g; %g  \n", d.P[0] - P0_[0], d.P[1] - P0_[1], d.P[2] - P0_[2]);
  std::printf("Lz = %g  dLz = %g \n", d.L[2], d.L[2] / L0_[2] - 1.0);
  std::printf("LL = %g  dLL = %g \n", LL, LL / LL0 - 1.0);
  std::printf("E = %g  dE = %g \n", E, E / E0_ - 1.0);

  const char* mode = it == 0 ? "w" : "a";
  auto open_csv = [&](const char* name, const char* header, File& fp) {
    fp.f = std::fopen(join(out_dir_, name).c_str(), mode);
    if (fp.f && it == 0) std::fprintf(fp.f, "%s\n", header);
    return fp.f != nullptr;
  };
  const std::string tt = ostream_double(t);
  {
    File fp;
    if (!open_csv("Lz(t).csv", "it, t, Lz, Lz_rel_err, Lz_rel_err_magnitude", fp)) return false;
    std::fprintf(fp.f, "%d, %s, %.15f, %s, %s\n", it_all, tt.c_str(), d.L[2],
                 ostream_double(d.L[2] / L0_[2] - 1.0).c_str(),
                 ostream_double(std::fabs(d.L[2] / L0_[2] - 1.0)).c_str());
  }
  {
    File fp;
    if (!open_csv("LL(t).csv", "it, t, LL, LL_rel_err, LL_rel_err_magnitude", fp)) return false;
    std::fprintf(fp.f, "%d, %s, %.15f, %s, %s\n", it_all, tt.c_str(), LL,
                 ostream_double(LL / LL0 - 1.0).c_str(),
                 ostream_double(std::fabs(LL / LL0 - 1.0)).c_str());
  }
  {
    File fp;
    if (!open_csv("Imp(t).csv", "it, t, Imp, Imp_abs_err, Imp_rel_err, Imp_rel_err_magnitude",
                  fp))
      return false;
    std::fprintf(fp.f, "%d, %s, %.15f, %s, %s, %s\n", it_all, tt.c_str(), Imp,
                 ostream_double(Imp - Imp0).c_str(), ostream_double(Imp / Imp0 - 1.0).c_str(),
                 ostream_double(std::fabs(Imp / Imp0 - 1.0)).c_str());
  }
  {
    File fp;
    if (!open_csv("E(t).csv", "it, t, E, E_rel_err, E_rel_err_magnitude, Ek, Ep", fp))
      return false;
    std::fprintf(fp.f, "%d, %s, %.15f, %s, %s, %s, %s\n", it_all, tt.c_str(), E,
                 ostream_double(E / E0_ - 1.0).c_str(),
                 ostream_double(std::fabs(E / E0_ - 1.0)).c_str(),
                 ostream_double(0.5 * d.Ek).c_str(), ostream_double(0.5 * d.Ep).c_str());
  }
  return true;
}

bool write_sample(const std::string& out_dir, int it, int it_all, double t, const double pos[3],
                  const double vel[3], const double acc[3], double phi) {
  File fp(join(out_dir, "sample.csv"), it == 1 ? "w" : "a");
  if (!fp.f) return false;
  if (it == 1) std::fprintf(fp.f, "it, t, i, r, fi, z, acc, phi, ax, ay, az\n");
  const double fi = std::atan2(pos[1], pos[0]);
  const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
  (void)vel;  // vr / vfi are computed but not written by the reference
  const double a = std::sqrt(acc[0] * acc[0] + acc[1] * acc[1] + acc[2] * acc[2]);
  std::fprintf(fp.f, "%d, %s, %d, %s, %s, %s, %s, %s, %s, %s, %s\n", it_all,
               ostream_double(t).c_str(), 1, ostream_double(r).c_str(),
               ostream_double(fi).c_str(), ostream_double(pos[2]).c_str(),
               ostream_double(a).c_str(), ostream_double(phi).c_str(),
               ostream_double(acc[0]).c_str(), ostream_double(acc[1]).c_str(),
               ostream_double(acc[2]).c_str());
  return true;
}

bool write_x_slice(const std::string& path, const char* header, int nx, double sim_l, double dx,
                   const std::vector<double>& cell_value, int decimals) {
  File fp(path, "w");
  if (!fp.f) return false;
  std::fprintf(fp.f, "%s\n", header);
  const int iy = nx / 2, iz = nx / 2;
  for (int i = 0; i < nx; ++i) {
    const double x = -sim_l + i * dx;
    char xs[64];
    auto res = std::to_chars(xs, xs + sizeof(xs) - 1, x);  // std::format "{}": shortest form
    *res.ptr = '\0';
    const size_t at = (size_t)i + (size_t)iy * nx + (size_t)iz * nx * nx;
    std::fprintf(fp.f, "%s, %.*f\n", xs, decimals, cell_value[at]);
  }
  return true;
}

bool write_cell_phi(const std::string& out_dir, int iter, int nx, double sim_l, double dx,
                    const std::vector<double>& cell_phi) {
  const std::string dir = join(out_dir, "phi");
  ::mkdir(dir.c_str(), 0755);
  return write_x_slice(join(dir, "phi_" + width5(iter) + ".csv"), "x, phi", nx, sim_l, dx,
                       cell_phi, 7);
}

bool write_index_list(const std::string& path, const std::vector<double>& values) {
  File fp(path, "w");
  if (!fp.f) return false;
  std::fprintf(fp.f, "i, acc\n");
  for (size_t i = 0; i < values.size(); ++i) std::fprintf(fp.f, "%zu, %.7f\n", i, values[i]);
  return true;
}

bool write_count_slice(const std::string& path, int nx, const std::vector<int>& count) {
  File fp(path, "w");
  if (!fp.f) return false;
  std::fprintf(fp.f, "xi, counts\n");
  const size_t row = (size_t)(nx / 2) * nx + (size_t)(nx / 2) * nx * nx;
  for (int i = 0; i < nx; ++i)
    std::fprintf(fp.f, "%d, %s\n", i, ostream_double((double)count[row + i]).c_str());
  return true;
}

bool write_count_distribution(const std::string& dir, const std::vector<int>& count) {
  std::map<int, long long> dist;
  for (int c : count) ++dist[c];
  File fp(join(dir, "particles_in_cell.txt"), "w");
  File fs(join(dir, "particles_in_cell_partial_sum.txt"), "w");
  if (!fp.f || !fs.f) return false;
  std::fprintf(fp.f, "n_particles, cells_with_so_many_particles\n");
  double partial = 0.;
  size_t k = 0;
  for (auto it = dist.rbegin(); it != dist.rend(); ++it, ++k) {
    std::fprintf(fp.f, "%d, %lld\n", it->first, it->second);
    partial += (double)it->first * (double)it->second;
    if (k > 0) std::fprintf(fs.f, "%zu, %s\n", k, ostream_double(partial).c_str());
  }
  return true;
}

}  // namespace khrr_host
