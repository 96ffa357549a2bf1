// khrr_nbody -- drop-in host driver for KHRR-NBody's direct-sum path.
//
// Replaces run() of nbody/src/nbody.cu (:1010-1711) on top of the C-ABI of
// include/khrr.h.  Inputs and outputs are the reference's files under a
// data directory laid out as data/nbody/{ini,bin,out}:
//   ini/__start_galaxies.ini, __start_nbody.ini, __gr_par.ini,
//   ini/start_S{k}.txt, start_DM{k}.txt                         (read)
//   bin/LIE_0.bin, bin/S_{:5}.bin, bin/DM_{:5}.bin                (written)
//   out/E(t).csv, Imp(t).csv, LL(t).csv, Lz(t).csv, sample.csv,
//   out/phi/phi_{:5}.csv                                        (written)
// Step order is the reference's: predictor -> ACCEL(x*) -> corrector
// (:1525-1582); every is_grav = round(dtsave/dtgrav) steps a save: PHI,
// binning on the 200^3 grid of dx 0.2 (:1134-1172), cell potential slice,
// sample row, conservation laws and, every 10th save, the snapshots.
//
// Deliberate differences (each selectable back where it is a choice):
//   * the force accumulator is overwritten every step; --compat-accumulate
//     reproduces the reference's never-cleared acct_ (nbody.cu:1560);
//   * conservation sums run as FP64 reductions on the device;
//   * restart (i_cont > 0) reads bin/S_{:5}.bin and bin/DM_{:5}.bin with
//     the names the writer uses (the reference builds a literal
//     "S_%5d.bin" and never reads DM, :1251,1274);
//   * the cell-potential buffer is accumulated across saves as in the
//     reference (never cleared, :1628-1634); --clear-cell-phi clears it.
//
// --debug-dumps writes the files of convert_particles (nbody.cu:287-701)
// to data/nbody/debug after the initial forces (debug_dumps below).
//
// --solver wave selects the reference's other force (NBodySolver::Wave,
// :1128-1143, hard-wired off there): calc_acceleration_by_grid (:183-285)
// bins the bodies, advances the damped-wave field (20000 iterations at
// setup, dtgrav/dt_wave per step, the field kept between calls), and reads
// -grad(phi) and phi at each body's cell.  The save block then skips PHI
// and the cell-potential sum and writes the wave field (:1601-1666).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <type_traits>
#include <vector>

#include "khrr.h"
#include "nbody_io.hpp"

using namespace khrr_host;

namespace {

struct Options {
  std::string data = "data/nbody";
  bool compat_accumulate = false;
  bool compensated = false;
  bool clear_cell_phi = false;
  long max_steps = -1;
  int grid_nx = 200;
  double grid_dx = 0.2;
  bool json = false;
  bool wave = false;
  int setup_iters = 20000;  // _iterations_to_setup, nbody.cu:1143
  bool debug_dumps = false;
  std::string tune_cache;  // autotuner cache (JSON) to load, or empty
};

void usage() {
  std::fprintf(stderr,
               "usage: khrr_nbody [--data DIR] [--compat-accumulate] [--compensated]\n"
               "                  [--clear-cell-phi] [--max-steps N] [--grid NX DX] [--json]\n"
               "                  [--solver nbody|wave] [--setup-iters N] [--debug-dumps]\n"
               "                  [--tune-cache FILE]\n");
}

bool parse(int argc, char** argv, Options& o) {
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--data" && i + 1 < argc) {
      o.data = argv[++i];
    } else if (a == "--compat-accumulate") {
      o.compat_accumulate = true;
    } else if (a == "--compensated") {
      o.compensated = true;
    } else if (a == "--clear-cell-phi") {
      o.clear_cell_phi = true;
    } else if (a == "--max-steps" && i + 1 < argc) {
      o.max_steps = std::atol(argv[++i]);
    } else if (a == "--grid" && i + 2 < argc) {
      o.grid_nx = std::atoi(argv[++i]);
      o.grid_dx = std::atof(argv[++i]);
    } else if (a == "--json") {
      o.json = true;
    } else if (a == "--solver" && i + 1 < argc) {
      const std::string v = argv[++i];
      if (v != "nbody" && v != "wave") {
        usage();
        return false;
      }
      o.wave = v == "wave";
    } else if (a == "--debug-dumps") {
      o.debug_dumps = true;
    } else if (a == "--tune-cache" && i + 1 < argc) {
      o.tune_cache = argv[++i];
    } else if (a == "--setup-iters" && i + 1 < argc) {
      o.setup_iters = std::atoi(argv[++i]);
      if (o.setup_iters < 0) return false;
    } else {
      usage();
      return false;
    }
  }
  return o.grid_nx > 0 && o.grid_dx > 0;
}

#define CHECK_HIP(call)                                                                \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      std::fprintf(stderr, "%s failed: %s\n", #call, hipGetErrorString(e_));           \
      return 1;                                                                        \
    }                                                                                  \
  } while (0)

#define CHECK_K(ctx, call)                                                          \
  do {                                                                              \
    int rc_ = (call);                                                               \
    if (rc_ != 0) {                                                                 \
      std::fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, khrr_last_error(ctx)); \
      return 1;                                                                     \
    }                                                                               \
  } while (0)

// Device buffers of one run, freed on scope exit.
struct DevBuffers {
  std::vector<void*> ptrs;
  template <typename T>
  T* alloc(size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, (count ? count : 1) * sizeof(T)) != hipSuccess) return nullptr;
    if (hipMemset(p, 0, (count ? count : 1) * sizeof(T)) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return static_cast<T*>(p);
  }
  ~DevBuffers() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

// Restart state from the snapshot pair of save `it` plus LIE_0.bin.
bool load_restart(const std::string& bin, int i_cont, const std::vector<GalaxyProperties>& gal,
                  Particles& p, double& t, double L0[3], double P0[3], double& E0,
                  std::string& err) {
  auto read_bin = [&](const std::string& name, int i0, int count) {
    const std::string path = bin + "/" + name + "_" + width5(i_cont) + ".bin";
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
      err = "can't open " + path;
      return false;
    }
    int cnt = 0;
    if (std::fread(&cnt, sizeof(int), 1, f) != 1 || std::fread(&t, sizeof(double), 1, f) != 1) {
      std::fclose(f);
      err = "short snapshot header in " + path;
      return false;
    }
    if (cnt != count) {
      std::fclose(f);
      err = path + " holds " + std::to_string(cnt) + " bodies, the galaxies have " +
            std::to_string(count);
      return false;
    }
    bool ok = true;
    const size_t N = (size_t)p.n;
    for (int i = i0; ok && i < i0 + count; ++i) {
      double r[6];
      ok = std::fread(r, sizeof(double), 6, f) == 6;
      for (int k = 0; ok && k < 3; ++k) {
        p.pos[k * N + i] = r[k];
        p.vel[k * N + i] = r[3 + k];
      }
    }
    std::fclose(f);
    if (!ok) err = "short snapshot " + path + " (fewer than " + std::to_string(count) + " bodies)";
    return ok;
  };
  int ns = 0, nd = 0;
  for (auto& g : gal) {
    ns += g.get_n_star_particles();
    nd += g.get_n_dark_particles();
  }
  if (ns > 0 && !read_bin("S", 0, ns)) return false;
  if (nd > 0 && !read_bin("DM", ns, nd)) return false;
  FILE* f = std::fopen((bin + "/LIE_0.bin").c_str(), "rb");
  if (!f) {
    err = "can't open " + bin + "/LIE_0.bin";
    return false;
  }
  bool ok = std::fread(L0, sizeof(double), 3, f) == 3 && std::fread(P0, sizeof(double), 3, f) == 3 &&
            std::fread(&E0, sizeof(double), 1, f) == 1;
  std::fclose(f);
  if (!ok) err = "short LIE_0.bin";
  return ok;
}

// Debug dumps of convert_particles (nbody.cu:287-701, with the checks of
// :35-127), written to data/nbody/debug at the place of its commented-out
// call after the initial forces (:1488-1504), from the current state: the
// FP64 positions and masses, the body potential phi and acceleration acc.
// need_init_phi is true there: the binning feeds a wave field started from
// rest, advanced `iters` iterations (the call names `_iterations`, which
// run() does not define; --setup-iters gives the count).  The reference
// sums with FP64 atomics; here every per-cell sum runs in body index order.
// The wave field and the cell acceleration are FP32 (the stencil's format);
// the bodies' cell potential and its gradient are FP64 as in the reference.
int debug_dumps(khrr_ctx* ctx, const std::string& dir, int n, int nx, double dx, double sim_l,
                const khrr_grid& grid, const khrr_wave& wave, int iters, const double* d_pos,
                const double* d_mass, const float* d_phi, const float* d_acc) {
  ::mkdir(dir.c_str(), 0755);
  const size_t N = (size_t)n;
  const long long G = (long long)nx * nx * nx;
  DevBuffers db;
  int* cell = db.alloc<int>(N);
  int* idc = db.alloc<int>(N);
  int* count = db.alloc<int>(G);
  int* off = db.alloc<int>(G);
  double* cmass = db.alloc<double>(G);
  float* w[3] = {db.alloc<float>(G), db.alloc<float>(G), db.alloc<float>(G)};
  float* src = db.alloc<float>(G);
  float* cacc = db.alloc<float>(3 * G);
  float* cabs = db.alloc<float>(G);
  double* cphi_part = db.alloc<double>(G);
  double* cacc64 = db.alloc<double>(3 * G);
  double* cabs64 = db.alloc<double>(G);
  double* cval = db.alloc<double>(G);
  double* pabs = db.alloc<double>(N);
  double* cpabs = db.alloc<double>(N);
  float* pacc = db.alloc<float>(3 * N);
  double* pavg = db.alloc<double>(N);
  if (!pavg) {
    std::fprintf(stderr, "Error: device allocation failed\n");
    return 1;
  }
  auto path = [&](const char* name) { return dir + "/" + name; };
  auto to_host = [&](const auto* d, size_t count_, auto& v) {
    v.resize(count_);
    return hipMemcpy(v.data(), d, count_ * sizeof(v[0]), hipMemcpyDeviceToHost) == hipSuccess;
  };
  auto cells_f64 = [&](const auto* d, std::vector<double>& v) {
    using T = std::remove_cv_t<std::remove_pointer_t<decltype(d)>>;
    std::vector<T> h;
    if (!to_host(d, (size_t)G, h)) return false;
    v.assign(h.begin(), h.end());
    return true;
  };
  auto sum = [](const std::vector<double>& v) {
    double s = 0.;
    for (double x : v) s += x;
    return s;
  };
  bool ok = true;
  auto slice = [&](const char* name, const char* header, const std::vector<double>& v, int dec) {
    ok = ok && write_x_slice(path(name), header, nx, sim_l, dx, v, dec);
  };

  // binning, wave field from rest, its acceleration (nbody.cu:333-380)
  CHECK_K(ctx, khrr_p2mesh_bin(ctx, &grid, n, d_pos, d_mass, cell, idc, count, off, nullptr,
                               nullptr, cmass));
  CHECK_K(ctx, khrr_wave_source(ctx, &wave, cmass, src));
  float *prev = w[0], *curr = w[1], *next = w[2];
  CHECK_K(ctx, khrr_wave_iterate(ctx, &wave, &prev, &curr, &next, src, iters));
  CHECK_K(ctx, khrr_mesh_acc(ctx, nx, dx, curr, cacc, cabs));

  std::vector<double> cell_mass, cell_phi, pmass;
  std::vector<int> cnt, cid;
  if (!cells_f64(cmass, cell_mass) || !cells_f64(curr, cell_phi) || !to_host(count, G, cnt) ||
      !to_host(cell, N, cid) || !to_host(d_mass, N, pmass)) {
    std::fprintf(stderr, "Error: debug dump copy failed\n");
    return 1;
  }
  // check_cell_particle_info_match (:48-66): host recount of the bodies
  {
    std::vector<int> manual(G, 0);
    for (int c : cid)
      if (c >= 0 && c < G) ++manual[c];
    for (long long c = 0; c < G; ++c)
      if (manual[c] != cnt[c])
        std::printf("cell info mismatch!!! %d : %d\n", manual[c], cnt[c]);
  }
  // check_energy (:35-46)
  double ep = 0.;
  for (long long c = 0; c < G; ++c) ep += 0.5 * cell_mass[c] * cell_phi[c];
  std::printf("ep: %g\n", ep);
  // check_particles_distribution (:68-96), check_mass (:98-127)
  ok = ok && write_count_distribution(dir, cnt);
  slice("mass_cell.txt", "x, mass", cell_mass, 6);
  std::printf("total mass in cells: %s\n", ostream_double(sum(cell_mass)).c_str());
  std::printf("total mass in particles: %s\n", ostream_double(sum(pmass)).c_str());
  // particle_counts, phi_cell (:412-446)
  ok = ok && write_count_slice(path("particle_counts.txt"), nx, cnt);
  slice("phi_cell.txt", "x, phi", cell_phi, 7);
  // phi_part: cell mean of the bodies' potential (:449-472)
  std::vector<double> h;
  CHECK_K(ctx, khrr_p2mesh_cell_phi(ctx, n, G, cell, idc, count, off, d_phi, cphi_part, 0));
  if (!cells_f64(cphi_part, h)) return 1;
  slice("phi_part.txt", "x, phi", h, 7);
  // acc_cell: |a| of the wave field (:475-488)
  if (!cells_f64(cabs, h)) return 1;
  slice("acc_cell.txt", "x, acc", h, 7);
  // acc_part: |a| of the gradient of the bodies' cell potential (:491-526)
  CHECK_K(ctx, khrr_mesh_acc_f64(ctx, nx, dx, cphi_part, cacc64, cabs64));
  if (!cells_f64(cabs64, h)) return 1;
  slice("acc_part.txt", "x, acc", h, 7);
  // nbody_cell_acc: mass-weighted cell mean of the bodies' |a| (:529-558)
  CHECK_K(ctx, khrr_acc_abs(ctx, n, d_acc, pabs));
  CHECK_K(ctx, khrr_p2mesh_cell_weighted(ctx, n, G, cell, idc, count, off, cmass, d_mass, pabs,
                                         cval));
  if (!cells_f64(cval, h)) return 1;
  slice("nbody_cell_acc.txt", "x, acc", h, 7);
  // acc_cell_cycle: cell acceleration to the bodies and back (:561-598)
  CHECK_K(ctx, khrr_mesh_to_particles(ctx, n, G, cell, cacc, nullptr, pacc, nullptr));
  CHECK_K(ctx, khrr_acc_abs(ctx, n, pacc, cpabs));
  CHECK_K(ctx, khrr_p2mesh_cell_weighted(ctx, n, G, cell, idc, count, off, cmass, d_mass, cpabs,
                                         cval));
  if (!cells_f64(cval, h)) return 1;
  slice("acc_cell_cycle.txt", "x, acc", h, 7);
  // cell_part_acc, nbody_part_acc (:602-645)
  if (!to_host(cpabs, N, h)) return 1;
  std::printf("cell part acc sum: %s\n", ostream_double(sum(h)).c_str());
  ok = ok && write_index_list(path("cell_part_acc.txt"), h);
  if (!to_host(pabs, N, h)) return 1;
  std::printf("nbody part acc sum: %s\n", ostream_double(sum(h)).c_str());
  ok = ok && write_index_list(path("nbody_part_acc.txt"), h);
  // mpart_from_cell, mpart (:648-697)
  CHECK_K(ctx, khrr_p2mesh_cell_avg(ctx, n, G, cell, count, cmass, pavg));
  if (!to_host(pavg, N, h)) return 1;
  std::printf("total mpart from cell: %s\n", ostream_double(sum(h)).c_str());
  ok = ok && write_index_list(path("mpart_from_cell.txt"), h);
  std::printf("total mpart: %s\n", ostream_double(sum(pmass)).c_str());
  ok = ok && write_index_list(path("mpart.txt"), pmass);
  if (!ok) {
    std::fprintf(stderr, "Error: can't write the debug dumps in %s\n", dir.c_str());
    return 1;
  }
  return 0;
}

int run(const Options& opt) {
  const std::string ini = opt.data + "/ini", bin = opt.data + "/bin", out = opt.data + "/out";
  std::string err;
  std::vector<GalaxyProperties> gal;
  StartInfo si;
  GravParams gp;
  if (!read_galaxy_properties(ini + "/__start_galaxies.ini", gal, err) || gal.empty() ||
      !read_start_info(ini + "/__start_nbody.ini", si, err) ||
      !read_gr_par(ini + "/__gr_par.ini", gp, err)) {
    std::fprintf(stderr, "Error: %s\n", err.empty() ? "No galaxy properties read" : err.c_str());
    return 1;
  }
  const double dtgrav = gp.dtgrav;
  const int is_grav = (int)(si.dtsave / dtgrav + 0.5);  // nbody.cu:1108
  if (is_grav < 1) {
    std::fprintf(stderr, "Error: dt_save / dtgrav rounds to %d saves\n", is_grav);
    return 1;
  }
  khrr_extpot ext;
  khrr_extpot_init(gp.Mh, gp.a, gp.Rh, gp.Mb, gp.b, gp.Rb, &ext);

  // particle state (nbody.cu:1247-1421)
  Particles p;
  {
    int ns = 0, nd = 0;
    for (auto& g : gal) {
      ns += g.n_star_particles;
      nd += g.n_dark_particles;
    }
    p.n = ns + nd;
    p.n_star = ns;
  }
  double t = 0.0, L0[3] = {0, 0, 0}, P0[3] = {0, 0, 0}, E0 = 0.0;
  int it = 1, itt = 1, itg = 1;
  double tgrav = dtgrav;
  if (si.i_cont > 0) {
    p.pos.assign(3 * (size_t)p.n, 0.0);
    p.vel.assign(3 * (size_t)p.n, 0.0);
    p.mass.assign(p.n, 0.0);
    p.eps2.assign(p.n, 0.0);
    int n0 = 0;
    for (auto& g : gal) {  // masses and softening per component, stars first
      for (int i = 0; i < g.get_n_star_particles(); ++i, ++n0) {
        p.mass[n0] = g.get_mass_star_particle();
        p.eps2[n0] = g.eps_stars * g.eps_stars;
      }
    }
    for (auto& g : gal) {
      for (int i = 0; i < g.get_n_dark_particles(); ++i, ++n0) {
        p.mass[n0] = g.get_mass_dark_particle();
        p.eps2[n0] = g.eps_dark * g.eps_dark;
      }
    }
    if (!load_restart(bin, si.i_cont, gal, p, t, L0, P0, E0, err)) {
      std::fprintf(stderr, "Error: %s\n", err.c_str());
      return 1;
    }
    it = (int)(t / si.dtsave) + 1;  // nbody.cu:1296,1310-1313
    itg = (int)(t / dtgrav) + 1;
    tgrav = t + dtgrav;
  } else if (!load_start_particles(ini, gal, p, err)) {
    std::fprintf(stderr, "Error: %s\n", err.c_str());
    return 1;
  }
  const int n = p.n;
  const size_t N = (size_t)n;
  std::printf("total particles: %d  star particles: %d\n", n, p.n_star);

  khrr_ctx* ctx = nullptr;
  if (khrr_ctx_create(0, &ctx) != 0) {
    std::fprintf(stderr, "Error: %s\n", khrr_last_error(nullptr));
    return 1;
  }
  struct CtxGuard {
    khrr_ctx* c;
    ~CtxGuard() { khrr_ctx_destroy(c); }
  } guard{ctx};
  // launch shapes tuned earlier on this kind of device (khrr_tune_save)
  if (!opt.tune_cache.empty() && khrr_tune_load(ctx, opt.tune_cache.c_str()) != 0) {
    std::fprintf(stderr, "Error: %s\n", khrr_last_error(ctx));
    return 1;
  }

  // host -> device (FP64 state; FP32 gravity inputs with h = eps^2/2)
  std::vector<float> pos32(3 * N), m32(N), h32(N);
  for (size_t i = 0; i < 3 * N; ++i) pos32[i] = (float)p.pos[i];
  for (size_t i = 0; i < N; ++i) {
    m32[i] = (float)p.mass[i];
    h32[i] = (float)(0.5 * p.eps2[i]);
  }
  const int nx = opt.grid_nx;
  const long long G = (long long)nx * nx * nx;
  const double sim_l = 0.5 * nx * opt.grid_dx;
  khrr_grid grid;
  for (int k = 0; k < 3; ++k) {
    grid.min[k] = -sim_l;
    grid.cell[k] = opt.grid_dx;
    grid.size[k] = nx;
  }
  DevBuffers db;
  double* d_pos = db.alloc<double>(3 * N);
  double* d_vel = db.alloc<double>(3 * N);
  double* d_post = db.alloc<double>(3 * N);
  double* d_velt = db.alloc<double>(3 * N);
  double* d_mass = db.alloc<double>(N);
  float* d_pos32 = db.alloc<float>(3 * N);
  float* d_m32 = db.alloc<float>(N);
  float* d_h32 = db.alloc<float>(N);
  float* d_acc = db.alloc<float>(3 * N);
  float* d_acct = db.alloc<float>(3 * N);
  float* d_phi = db.alloc<float>(N);
  int* d_cell = db.alloc<int>(N);
  int* d_idc = db.alloc<int>(N);
  int* d_count = db.alloc<int>(G);
  int* d_off = db.alloc<int>(G);
  double* d_cmass = db.alloc<double>(G);
  double* d_cphi = db.alloc<double>(G);
  // wave solver: three fields, the source and the cell acceleration
  float* d_wave[3] = {nullptr, nullptr, nullptr};
  float* d_src = nullptr;
  float* d_cacc = nullptr;
  if (opt.wave) {
    for (auto& f : d_wave) f = db.alloc<float>(G);
    d_src = db.alloc<float>(G);
    d_cacc = db.alloc<float>(3 * G);
    if (!d_cacc) {
      std::fprintf(stderr, "Error: device allocation failed\n");
      return 1;
    }
    for (auto& f : d_wave) CHECK_HIP(hipMemset(f, 0, G * sizeof(float)));
  }
  if (!d_cphi) {
    std::fprintf(stderr, "Error: device allocation failed\n");
    return 1;
  }
  CHECK_HIP(hipMemcpy(d_pos, p.pos.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice));
  CHECK_HIP(hipMemcpy(d_vel, p.vel.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice));
  CHECK_HIP(hipMemcpy(d_mass, p.mass.data(), N * sizeof(double), hipMemcpyHostToDevice));
  CHECK_HIP(hipMemcpy(d_pos32, pos32.data(), 3 * N * sizeof(float), hipMemcpyHostToDevice));
  CHECK_HIP(hipMemcpy(d_m32, m32.data(), N * sizeof(float), hipMemcpyHostToDevice));
  CHECK_HIP(hipMemcpy(d_h32, h32.data(), N * sizeof(float), hipMemcpyHostToDevice));

  const unsigned gflags = (opt.compensated ? KHRR_COMPENSATED : 0u);
  const unsigned step_flags = gflags | (opt.compat_accumulate ? KHRR_ACCUMULATE : 0u);
  ResultWriter writer(bin, out);
  if (si.i_cont > 0) writer.set_initial(L0, P0, E0);

  auto diag = [&](Diag& d) {
    double v[KHRR_DIAG_COUNT];
    if (khrr_diag(ctx, n, d_pos, d_vel, d_mass, d_phi, v) != 0) return false;
    d.Ek = v[KHRR_DIAG_EK];
    d.Ep = v[KHRR_DIAG_EP];
    for (int k = 0; k < 3; ++k) {
      d.P[k] = v[KHRR_DIAG_PX + k];
      d.L[k] = v[KHRR_DIAG_LX + k];
    }
    d.Vr_max = v[KHRR_DIAG_VR_MAX];
    d.Vfi_max = v[KHRR_DIAG_VFI_MAX];
    d.Vz_max = v[KHRR_DIAG_VZ_MAX];
    d.R_max = v[KHRR_DIAG_R_MAX];
    d.Z_max = v[KHRR_DIAG_Z_MAX];
    return true;
  };

  // wave constants (nbody.cu:1134-1143) on the driver's grid
  khrr_wave wave{};
  wave.nx = nx;
  wave.dx = opt.grid_dx;
  wave.c = 4574.337022617616;
  wave.dt = 0.5 * opt.grid_dx / (wave.c * std::sqrt(3.0));
  wave.diss_base = 0.1;
  wave.diss_extra = 0.01;
  wave.sim_l = sim_l;
  wave.bc_l = sim_l / 10.;
  float* w_prev = d_wave[0];
  float* w_curr = d_wave[1];
  float* w_next = d_wave[2];
  // calc_acceleration_by_grid (nbody.cu:183-285) at the FP64 positions x
  auto mesh_force = [&](const double* x, int iters, float* acc_out) {
    int rc = khrr_p2mesh_bin(ctx, &grid, n, x, d_mass, d_cell, d_idc, d_count, d_off, nullptr,
                             nullptr, d_cmass);
    if (!rc) rc = khrr_wave_source(ctx, &wave, d_cmass, d_src);
    if (!rc) rc = khrr_wave_iterate(ctx, &wave, &w_prev, &w_curr, &w_next, d_src, iters);
    if (!rc) rc = khrr_mesh_acc(ctx, nx, opt.grid_dx, w_curr, d_cacc, nullptr);
    if (!rc) rc = khrr_mesh_to_particles(ctx, n, G, d_cell, d_cacc, w_curr, acc_out, d_phi);
    return rc;
  };
  auto write_wave_field = [&](int frame) {
    std::vector<float> f(G);
    if (hipMemcpy(f.data(), w_curr, G * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      return false;
    write_cell_phi(out, frame, nx, sim_l, opt.grid_dx, std::vector<double>(f.begin(), f.end()));
    return true;
  };
  const int step_iters = (int)(dtgrav / wave.dt);  // nbody.cu:1556

  // initial PHI and ACCEL (nbody.cu:1449-1481)
  if (opt.wave) {
    CHECK_K(ctx, mesh_force(d_pos, opt.setup_iters, d_acc));
    if (!write_wave_field(0)) {
      std::fprintf(stderr, "Error: wave field copy failed\n");
      return 1;
    }
  } else {
    CHECK_K(ctx, khrr_accel(ctx, n, d_pos32, d_h32, n, d_pos32, d_m32, d_h32, d_acc, d_phi, gflags,
                            nullptr));
  }
  if (opt.debug_dumps) {
    if (debug_dumps(ctx, opt.data + "/debug", n, nx, opt.grid_dx, sim_l, grid, wave,
                    opt.setup_iters, d_pos, d_mass, d_phi, d_acc) != 0)
      return 1;
  }
  if (it == 1) {
    Diag d0;
    if (!diag(d0)) {
      std::fprintf(stderr, "Error: %s\n", khrr_last_error(ctx));
      return 1;
    }
    writer.write(d0, 0, 0.0, 0);
  }

  auto wall0 = std::chrono::steady_clock::now();
  long steps = 0;
  int it_grav = 0;
  double pair_evals = opt.wave ? 0.0 : (double)n * n;  // initial force
  do {
    // predictor (tn + dtgrav), gravity at x*, corrector (nbody.cu:1525-1582)
    CHECK_K(ctx, khrr_integrate(ctx, n, 0, dtgrav, tgrav - dtgrav, &ext, d_acc, d_post, d_velt,
                                d_pos, d_vel, nullptr, d_pos32));
    if (opt.wave) {
      CHECK_K(ctx, mesh_force(d_post, step_iters, d_acct));
    } else {
      CHECK_K(ctx, khrr_accel(ctx, n, d_pos32, d_h32, n, d_pos32, d_m32, d_h32, d_acct, nullptr,
                              step_flags, nullptr));
      pair_evals += (double)n * n;
    }
    CHECK_K(ctx, khrr_integrate(ctx, n, 1, dtgrav, tgrav, &ext, d_acct, d_pos, d_vel, d_post,
                                d_velt, d_acc, nullptr));
    ++steps;
    ++itg;
    t = tgrav;
    tgrav = itg * dtgrav;
    ++it_grav;
    if (it_grav >= is_grav) {
      // save block (nbody.cu:1597-1707)
      if (opt.wave) {
        // phi is the mesh potential at the bodies; the wave field is written
        if (!write_wave_field(itt * it)) {
          std::fprintf(stderr, "Error: wave field copy failed\n");
          return 1;
        }
      } else {
        CHECK_K(ctx, khrr_accel(ctx, n, d_pos32, d_h32, n, d_pos32, d_m32, d_h32, nullptr, d_phi,
                                gflags, nullptr));
        pair_evals += (double)n * n;
        CHECK_K(ctx, khrr_p2mesh_bin(ctx, &grid, n, d_pos, d_mass, d_cell, d_idc, d_count, d_off,
                                     nullptr, nullptr, d_cmass));
        CHECK_K(ctx, khrr_p2mesh_cell_phi(ctx, n, G, d_cell, d_idc, d_count, d_off, d_phi, d_cphi,
                                          opt.clear_cell_phi ? 0 : 1));
        std::vector<double> cphi(G);
        CHECK_HIP(hipMemcpy(cphi.data(), d_cphi, G * sizeof(double), hipMemcpyDeviceToHost));
        write_cell_phi(out, itt * it, nx, sim_l, opt.grid_dx, cphi);
      }
      if (n > 1) {
        double p1[3], v1[3], a1[3];
        float a1f[3], phi1 = 0.f;
        for (int k = 0; k < 3; ++k) {
          CHECK_HIP(hipMemcpy(&p1[k], d_pos + k * N + 1, sizeof(double), hipMemcpyDeviceToHost));
          CHECK_HIP(hipMemcpy(&v1[k], d_vel + k * N + 1, sizeof(double), hipMemcpyDeviceToHost));
          CHECK_HIP(hipMemcpy(&a1f[k], d_acc + k * N + 1, sizeof(float), hipMemcpyDeviceToHost));
          a1[k] = a1f[k];
        }
        CHECK_HIP(hipMemcpy(&phi1, d_phi + 1, sizeof(float), hipMemcpyDeviceToHost));
        write_sample(out, it, it * itt, t, p1, v1, a1, phi1);
      }
      if (it > 1 && it % 10 == 0) {
        CHECK_HIP(hipMemcpy(p.pos.data(), d_pos, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
        CHECK_HIP(hipMemcpy(p.vel.data(), d_vel, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
        write_particles_bin(bin, "S", 0, p.n_star, p, it, t);
        if (n > p.n_star) write_particles_bin(bin, "DM", p.n_star, n - p.n_star, p, it, t);
      }
      Diag d;
      if (!diag(d)) {
        std::fprintf(stderr, "Error: %s\n", khrr_last_error(ctx));
        return 1;
      }
      writer.write(d, it, t, it * itt);
      auto now = std::chrono::steady_clock::now();
      std::printf("<Time_frame> = %.3f s, frame = %d, iter = %d\n",
                  std::chrono::duration<double>(now - wall0).count(), it, itt);
      ++it;
      itt = 0;
      it_grav = 0;
    }
    ++itt;
  } while (t < si.tmax && (opt.max_steps < 0 || steps < opt.max_steps));
  CHECK_K(ctx, khrr_ctx_sync(ctx));
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
  if (opt.json)
    std::printf("{\"n\": %d, \"steps\": %ld, \"t\": %.9g, \"wall_s\": %.6f, "
                "\"pair_evals\": %.6e, \"Gpairs_per_s\": %.3f}\n",
                n, steps, t, wall, pair_evals, pair_evals / wall / 1e9);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  Options opt;
  if (!parse(argc, argv, opt)) return 2;
  return run(opt);
}
