/*
 * khrr_oracle.c -- CPU FP64 restatement of the KHRR-NBody hot path.
 * TEST INFRASTRUCTURE ONLY (see khrr_oracle.h): the checker for the HIP
 * path and the timed CPU baseline, never part of the product.
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include "khrr_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* common.cuh:7 -- the reference's truncated pi */
#define ORC_PI 3.14159265358979
/* common.cuh:8 */
#define ORC_BLOCK 512

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* nbody.cu:1111-1120: host constants of the halo + bulge potential.      */
void orc_extpot_init(double Mh, double a, double Rh, double Mb, double b, double Rb,
                     orc_extpot* e) {
  const double Rh2 = 3.0 * Rh;
  const double rcore1 = Rh / a;
  const double rbcore1 = 1.0 / b;
  const double rbcore2 = rbcore1 * rbcore1;
  const double root1 = sqrt(1.0 + (Rb * Rb) * rbcore2);
  const double con = Mh / (rcore1 - atan(rcore1));
  const double const1 = Mb / (b * log(Rb * rbcore1 + root1) - Rb / root1);
  const double c_phi_h =
      con / a * (0.5 * log(Rh2 * Rh2 / a / a + 1.0) + atan(Rh2 / a) * a / Rh2) + Mh / Rh2;
  const double c_phi_b = Mb / Rb - const1 * log(Rb / b + root1) / Rb;
  const double Mh_inf = Mh * (Rh2 / a - atan(Rh2 / a)) / (rcore1 - atan(rcore1));
  e->Mh = Mh;
  e->Mh_inf = Mh_inf;
  e->a = a;
  e->Rh = Rh;
  e->Mb = Mb;
  e->b = b;
  e->Rb = Rb;
  e->Rh2 = Rh2;
  e->con = con;
  e->const1 = const1;
  e->c_phi_h = c_phi_h;
  e->c_phi_b = c_phi_b;
}

/* nbody-kernel.cuh:8-40 */
void orc_dev_fex(const orc_extpot* e, const double p[3], double f[3]) {
  const double rr = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  if (rr > 0.0) {
    const double rr3 = rr * rr * rr;
    const double rrcore = rr / e->a;
    const double rrbcore1 = rr / e->b;
    const double root1 = sqrt(1.0 + rrbcore1 * rrbcore1);
    double forcehalo, forceblg1;
    if (rr < e->Rh2)
      forcehalo = -e->con * (rrcore - atan(rrcore)) / rr3;
    else
      forcehalo = -e->Mh_inf / rr3;
    if (rr < e->Rb)
      forceblg1 = -e->const1 * (e->b * log(rrbcore1 + root1) - rr / root1) / rr3;
    else
      forceblg1 = -e->Mb / rr3;
    const double forcesph = forceblg1 + forcehalo;
    f[0] = forcesph * p[0];
    f[1] = forcesph * p[1];
    f[2] = forcesph * p[2];
  } else {
    f[0] = f[1] = f[2] = 0.0;
  }
}

/* ---------------------------------------------------------------------- */
/* One i row of ACCEL_kernel (nbody-kernel.cuh:108-141) and PHI_kernel
 * (:55-85): j ascending, dp = r_j - r_i, denominator = sqrt(|dp|^2 +
 * (h_i + h_j)), k = m / denominator^3. */
static void accel_row(double xi, double yi, double zi, double hi, int n_j, const double* pos_j,
                      const double* m_j, const double* h_j, double* fx, double* fy, double* fz,
                      double* phis) {
  const double* xj = pos_j;
  const double* yj = pos_j + n_j;
  const double* zj = pos_j + 2 * (size_t)n_j;
  double sx = 0.0, sy = 0.0, sz = 0.0, sp = 0.0;
  for (int j = 0; j < n_j; ++j) {
    const double dx = xj[j] - xi;
    const double dy = yj[j] - yi;
    const double dz = zj[j] - zi;
    const double den = sqrt((dx * dx + dy * dy + dz * dz) + (hi + h_j[j]));
    const double k = m_j[j] / (den * den * den);
    sx += dx * k;
    sy += dy * k;
    sz += dz * k;
    sp += m_j[j] / den;
  }
  *fx = sx;
  *fy = sy;
  *fz = sz;
  *phis = sp;
}

void orc_accel(int n_i, const double* pos_i, const double* h_i, int n_j, const double* pos_j,
               const double* m_j, const double* h_j, double* acc, double* phi, int accumulate) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n_i; ++i) {
    double fx, fy, fz, ps;
    accel_row(pos_i[i], pos_i[n_i + i], pos_i[2 * (size_t)n_i + i], h_i[i], n_j, pos_j, m_j, h_j,
              &fx, &fy, &fz, &ps);
    if (acc) {
      if (accumulate) {
        acc[i] = acc[i] + fx;
        acc[n_i + i] = acc[n_i + i] + fy;
        acc[2 * (size_t)n_i + i] = acc[2 * (size_t)n_i + i] + fz;
      } else {
        acc[i] = fx;
        acc[n_i + i] = fy;
        acc[2 * (size_t)n_i + i] = fz;
      }
    }
    if (phi) phi[i] = accumulate ? phi[i] - ps : -ps;
  }
}

void orc_accel_rows(int n_rows, const int* rows, int n_i, const double* pos_i, const double* h_i,
                    int n_j, const double* pos_j, const double* m_j, const double* h_j,
                    double* acc, double* phi) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < n_rows; ++r) {
    const int i = rows[r];
    double fx, fy, fz, ps;
    accel_row(pos_i[i], pos_i[n_i + i], pos_i[2 * (size_t)n_i + i], h_i[i], n_j, pos_j, m_j, h_j,
              &fx, &fy, &fz, &ps);
    if (acc) {
      acc[r] = fx;
      acc[n_rows + r] = fy;
      acc[2 * (size_t)n_rows + r] = fz;
    }
    if (phi) phi[r] = -ps;
  }
}

/* ---------------------------------------------------------------------- */
/* kernelNbody_integTime nbody-kernel.cuh:144-194 */
void orc_integrate(int n, int istep, double dt, double t, const orc_extpot* ext,
                   const double* acc, double* pos_t, double* vel_t, const double* pos,
                   const double* vel, double* acc0) {
  (void)t;
  const size_t N = (size_t)n;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double r[3] = {pos[i], pos[N + i], pos[2 * N + i]};
    double v[3] = {vel[i], vel[N + i], vel[2 * N + i]};
    double fex[3] = {0.0, 0.0, 0.0};
    if (ext) orc_dev_fex(ext, r, fex);
    double f[3];
    for (int k = 0; k < 3; ++k) f[k] = acc[k * N + i] + fex[k];
    if (istep == 0) {
      double vt[3];
      for (int k = 0; k < 3; ++k) vt[k] = v[k] + dt * f[k];
      for (int k = 0; k < 3; ++k) pos_t[k * N + i] = r[k] + 0.5 * dt * (v[k] + vt[k]);
      for (int k = 0; k < 3; ++k) vel_t[k * N + i] = vt[k];
    } else {
      double vt[3], vtt[3];
      for (int k = 0; k < 3; ++k) vt[k] = vel_t[k * N + i];
      for (int k = 0; k < 3; ++k) vtt[k] = 0.5 * (vt[k] + v[k] + dt * f[k]);
      for (int k = 0; k < 3; ++k) pos_t[k * N + i] = r[k];
      for (int k = 0; k < 3; ++k) vel_t[k * N + i] = vtt[k];
      if (acc0)
        for (int k = 0; k < 3; ++k) acc0[k * N + i] = acc[k * N + i];
    }
  }
}

/* ---------------------------------------------------------------------- */
/* result() nbody.cu:797-833, serial FP64 in index order. */
void orc_diag(int n, const double* pos, const double* vel, const double* mass, const double* phi,
              double* out) {
  const size_t N = (size_t)n;
  double Ek = 0, Ep = 0, Px = 0, Py = 0, Pz = 0, Lx = 0, Ly = 0, Lz = 0;
  double M = 0, Mx = 0, My = 0, Mz = 0;
  double Vr_max = 0, Vfi_max = 0, Vz_max = 0, R_max = 0, Z_max = 0;
  for (size_t i = 0; i < N; ++i) {
    const double x = pos[i], y = pos[N + i], z = pos[2 * N + i];
    const double vx = vel[i], vy = vel[N + i], vz = vel[2 * N + i];
    const double r = sqrt(x * x + y * y);
    double vr, vfi;
    if (r > 0.0) {
      vr = (vx * x + vy * y) / r;
      vfi = (vy * x - vx * y) / r;
    } else {
      vr = 0.0;
      vfi = 0.0;
    }
    const double m = mass[i];
    Lx += (vz * y - vy * z) * m;
    Ly += (vx * z - vz * x) * m;
    Lz += (vy * x - vx * y) * m;
    Px += m * vx;
    Py += m * vy;
    Pz += m * vz;
    Ek += m * (vx * vx + vy * vy + vz * vz);
    if (phi) Ep += m * phi[i];
    M += m;
    Mx += m * x;
    My += m * y;
    Mz += m * z;
    Vr_max = fmax(Vr_max, fabs(vr));
    Vfi_max = fmax(Vfi_max, vfi);
    Vz_max = fmax(Vz_max, fabs(vz));
    R_max = fmax(R_max, r);
    Z_max = fmax(Z_max, fabs(z));
  }
  const double v[17] = {Ek, Ep, Px, Py, Pz, Lx, Ly, Lz, M, Mx, My, Mz,
                        Vr_max, Vfi_max, Vz_max, R_max, Z_max};
  memcpy(out, v, sizeof(v));
}

/* ---------------------------------------------------------------------- */
/* nbody.cu:1468-1594: initial PHI+ACCEL, then per step predictor ->
 * ACCEL(x*) -> corrector. */
void orc_run(int n, double* pos, double* vel, const double* mass, const double* h, double dt,
             int steps, const orc_extpot* ext, int compat_accumulate, double* diag_out,
             int diag_every) {
  if (diag_every < 1) diag_every = 1;
  const size_t N = (size_t)n;
  double* acc = (double*)calloc(3 * N, sizeof(double));
  double* acct = (double*)calloc(3 * N, sizeof(double));
  double* post = (double*)calloc(3 * N, sizeof(double));
  double* velt = (double*)calloc(3 * N, sizeof(double));
  double* phi = (double*)calloc(N, sizeof(double));
  orc_accel(n, pos, h, n, pos, mass, h, acc, diag_out ? phi : NULL, 0);
  if (diag_out) orc_diag(n, pos, vel, mass, phi, diag_out);
  double tgrav = dt;
  for (int s = 0; s < steps; ++s) {
    orc_integrate(n, 0, dt, tgrav - dt, ext, acc, post, velt, pos, vel, acc);
    orc_accel(n, post, h, n, post, mass, h, acct, NULL, compat_accumulate);
    orc_integrate(n, 1, dt, tgrav, ext, acct, pos, vel, post, velt, acc);
    tgrav = (s + 2) * dt;
    if (diag_out && (s + 1) % diag_every == 0) {
      orc_accel(n, pos, h, n, pos, mass, h, NULL, phi, 0);
      orc_diag(n, pos, vel, mass, phi, diag_out + (size_t)((s + 1) / diag_every) * 17);
    }
  }
  free(acc);
  free(acct);
  free(post);
  free(velt);
  free(phi);
}

/* ---------------------------------------------------------------------- */
int orc_sat_trunc(double v) {
  if (v != v) return 0;
  if (v >= 2147483648.0) return INT_MAX;
  if (v <= -2147483649.0) return INT_MIN;
  return (int)v;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* p2mesh-kernel.cuh:11-57 (cell index), :62-123 (offsets), :126-137 (mass) */
int orc_p2mesh_bin(const double gmin[3], const double gcell[3], const int gsize[3], int n,
                   const double* pos, const double* mass, int* cell_id, int* id_in_cell,
                   int* count, int* offset, int* start_id_ref, int* block_incl,
                   double* cell_mass) {
  const long long G = (long long)gsize[0] * gsize[1] * gsize[2];
  if (gsize[0] <= 0 || gsize[1] <= 0 || gsize[2] <= 0 || G > INT_MAX) return -1;
  /* the index formula below (z clamped by size[1], y stride size[2]) must
   * stay inside the grid for every clamped (ix, iy, iz) */
  if ((long long)(gsize[1] - 1) * gsize[0] * gsize[1] + (long long)(gsize[1] - 1) * gsize[2] +
          (gsize[0] - 1) >= G)
    return -1;
  const size_t N = (size_t)n;
  const int sx = gsize[0], sy = gsize[1], sz = gsize[2];
  int* cid = (int*)malloc((N ? N : 1) * sizeof(int));
  for (long long c = 0; c < G; ++c) count[c] = 0;
  if (cell_mass)
    for (long long c = 0; c < G; ++c) cell_mass[c] = 0.0;
  for (size_t i = 0; i < N; ++i) {
    const int ix = clampi(orc_sat_trunc((pos[i] - gmin[0]) / gcell[0]), 0, gsize[0] - 1);
    const int iy = clampi(orc_sat_trunc((pos[N + i] - gmin[1]) / gcell[1]), 0, gsize[1] - 1);
    /* z clamped with gridSize.y, p2mesh-kernel.cuh:36 */
    const int iz = clampi(orc_sat_trunc((pos[2 * N + i] - gmin[2]) / gcell[2]), 0, gsize[1] - 1);
    /* y stride written `sz`, p2mesh-kernel.cuh:43 */
    const int c = iz * sx * sy + iy * sz + ix;
    cid[i] = c;
    if (cell_id) cell_id[i] = c;
    /* index-order rank == a sequential atomic counter */
    const int rank = count[c]++;
    if (id_in_cell) id_in_cell[i] = rank;
    if (cell_mass) cell_mass[c] += mass[i];
  }
  const long long B = (G + ORC_BLOCK - 1) / ORC_BLOCK;
  long long run = 0;
  for (long long b = 0; b < B; ++b) {
    const long long blk_start = run;
    const long long c1 = (b + 1) * ORC_BLOCK < G ? (b + 1) * ORC_BLOCK : G;
    for (long long c = b * ORC_BLOCK; c < c1; ++c) {
      if (offset) offset[c] = (int)run;
      if (start_id_ref) start_id_ref[c] = (int)blk_start;
      run += count[c];
    }
    if (block_incl) block_incl[b] = (int)run;
  }
  free(cid);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* wave_diss_iteration wave-kernel.cuh:5-103, macros expanded in order. */
void orc_wave_step(int nx, double dx, double c, double dt, double diss_base, double diss_extra,
                   double sim_l, double bc_l, double* phi_new, const double* phi,
                   const double* phi_old, const double* mass) {
  const int NX = nx;
  const double DX = dx, C = c, DT = dt;
#define ORC_AT(x, y, z) ((size_t)(x) + (size_t)(y) * NX + (size_t)(z) * NX * NX)
#pragma omp parallel for schedule(static)
  for (int iz = 0; iz < NX; ++iz) {
    for (int iy = 0; iy < NX; ++iy) {
      for (int ix = 0; ix < NX; ++ix) {
        const double x = ix * DX, y = iy * DX, z = iz * DX;
        const double rho = mass[ORC_AT(ix, iy, iz)] / (DX * DX * DX);
        const double G = 1.;
        const double f = 4 * ORC_PI * G * rho;
        double diss = diss_base;
        if (x > -(sim_l - bc_l)) {
          const double xr = sim_l - bc_l;
          diss += (diss_extra * (fabs(x - xr)) * (fabs(x - xr)));
        } else if (x < -(sim_l - bc_l)) {
          const double xl = -(sim_l - bc_l);
          diss += (diss_extra * (fabs(x - xl)) * (fabs(x - xl)));
        }
        if (y > (sim_l - bc_l)) {
          const double yt = sim_l - bc_l;
          diss += (diss_extra * (fabs(y - yt)) * (fabs(y - yt)));
        } else if (y < -(sim_l - bc_l)) {
          const double yb = -(sim_l - bc_l);
          diss += (diss_extra * (fabs(y - yb)) * (fabs(y - yb)));
        }
        if (z > (sim_l - bc_l)) {
          const double zf = sim_l - bc_l;
          diss += (diss_extra * (fabs(z - zf)) * (fabs(z - zf)));
        } else if (z < -(sim_l - bc_l)) {
          const double zn = -(sim_l - bc_l);
          diss += (diss_extra * (fabs(z - zn)) * (fabs(z - zn)));
        }
        const double Q = (diss * C * DT * 0.5);
        const double W = (1 + Q);
        const double K = ((C * C) * (DT * DT) / (DX * DX));
        const double K_MAIN = (K / W);
        const double K_F = (-(C * C) * (DT * DT) / W);
        const double K_ACTUAL = (2 * (1. - 3 * K) / W);
        const double K_OLD = ((Q - 1.) / W);
        const size_t at = ORC_AT(ix, iy, iz);
        if (ix == 0 || ix == NX - 1 || iy == 0 || iy == NX - 1 || iz == 0 || iz == NX - 1) {
          phi_new[at] = 0.;
        } else {
          phi_new[at] = phi[ORC_AT(ix - 1, iy, iz)] * K_MAIN + phi[ORC_AT(ix + 1, iy, iz)] * K_MAIN +
                        phi[ORC_AT(ix, iy - 1, iz)] * K_MAIN + phi[ORC_AT(ix, iy + 1, iz)] * K_MAIN +
                        phi[ORC_AT(ix, iy, iz - 1)] * K_MAIN + phi[ORC_AT(ix, iy, iz + 1)] * K_MAIN +
                        phi[at] * K_ACTUAL + phi_old[at] * K_OLD + f * K_F;
        }
      }
    }
  }
#undef ORC_AT
}

void orc_wave_iterate(int nx, double dx, double c, double dt, double diss_base,
                      double diss_extra, double sim_l, double bc_l, const double* prev0,
                      const double* curr0, const double* mass, int iters, double* phi_out,
                      double* prev_out) {
  const size_t G = (size_t)nx * nx * nx;
  double* b[3];
  for (int k = 0; k < 3; ++k) b[k] = (double*)malloc(G * sizeof(double));
  memcpy(b[0], prev0, G * sizeof(double));
  memcpy(b[1], curr0, G * sizeof(double));
  memset(b[2], 0, G * sizeof(double));
  double *prev = b[0], *curr = b[1], *next = b[2];
  for (int it = 0; it < iters; ++it) {
    orc_wave_step(nx, dx, c, dt, diss_base, diss_extra, sim_l, bc_l, next, curr, prev, mass);
    double* t = prev; /* swap(prev, curr); swap(curr, next) */
    prev = curr;
    curr = next;
    next = t;
  }
  memcpy(phi_out, curr, G * sizeof(double));
  if (prev_out) memcpy(prev_out, prev, G * sizeof(double));
  for (int k = 0; k < 3; ++k) free(b[k]);
}

/* ---------------------------------------------------------------------- */
/* computeCellPhiUnsorted p2mesh-kernel.cuh:152-167, particles in index
 * order: cell_phi[c] += phi[i] / count[c]. */
void orc_p2mesh_cell_phi(int n, long long G, const int* cell_id, const int* count,
                         const double* phi, double* cell_phi, int accumulate) {
  if (!accumulate)
    for (long long c = 0; c < G; ++c) cell_phi[c] = 0.0;
  for (int i = 0; i < n; ++i) {
    const int c = cell_id[i];
    cell_phi[c] += phi[i] / count[c];
  }
}

/* apply_particle_to_cell_weighted p2mesh-kernel.cuh:273-287: every body
 * adds its value times its share of the cell's mass, weight =
 * particle_mass / cell_mass, here in index order into a zeroed buffer. */
void orc_p2mesh_cell_weighted(int n, long long G, const int* cell_id, const double* cell_mass,
                              const double* mass, const double* value, double* cell_value) {
  for (long long c = 0; c < G; ++c) cell_value[c] = 0.0;
  for (int i = 0; i < n; ++i) {
    const int c = cell_id[i];
    const double weight = mass[i] / cell_mass[c];
    cell_value[c] += value[i] * weight;
  }
}

/* apply_cell_to_particles_avg p2mesh-kernel.cuh:256-271. */
void orc_p2mesh_cell_avg(int n, const int* cell_id, const int* count, const double* cell_value,
                         double* out) {
  for (int i = 0; i < n; ++i) {
    const int c = cell_id[i];
    out[i] = count[c] > 0 ? cell_value[c] / count[c] : -1.;
  }
}

/* acc_abs p2mesh-kernel.cuh:289-298: norm3 = sqrt(dot3(v, v)) with dot3
 * summed x, y, z left to right (common.cuh:84-96). */
void orc_acc_abs(int n, const double* v, double* out) {
  const size_t N = (size_t)n;
  for (int i = 0; i < n; ++i) {
    const double x = v[i], y = v[N + i], z = v[2 * N + i];
    out[i] = sqrt(x * x + y * y + z * z);
  }
}

/* ---------------------------------------------------------------------- */
/* acc_field p2mesh-kernel.cuh:169-224 with every cell at its own (ix, iy,
 * iz) (the 3-D stencil the kernel's indexing describes): per axis a
 * one-sided difference on the low face, on the high face, else half the
 * central difference; then times -1/dx.  acc: vec3 planes of nx^3;
 * acc_abs (may be NULL): sqrt(ax*ax + ay*ay + az*az) (norm3, common.cuh:94). */
static double orc_diff1(int i, int nx, double minus, double c, double plus) {
  if (i == 0) return plus - c;
  if (i == nx - 1) return c - minus;
  return 0.5 * (plus - minus);
}

void orc_mesh_acc(int nx, double dx, const double* phi, double* acc, double* acc_abs) {
  const size_t N = (size_t)nx, G = N * N * N;
  const double coef = -1. / dx;
  for (int iz = 0; iz < nx; ++iz)
    for (int iy = 0; iy < nx; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const size_t at = (size_t)ix + (size_t)iy * N + (size_t)iz * N * N;
        const double c = phi[at];
        double gx = orc_diff1(ix, nx, ix > 0 ? phi[at - 1] : 0.0, c, ix + 1 < nx ? phi[at + 1] : 0.0);
        double gy = orc_diff1(iy, nx, iy > 0 ? phi[at - N] : 0.0, c, iy + 1 < nx ? phi[at + N] : 0.0);
        double gz = orc_diff1(iz, nx, iz > 0 ? phi[at - N * N] : 0.0, c,
                              iz + 1 < nx ? phi[at + N * N] : 0.0);
        gx *= coef;
        gy *= coef;
        gz *= coef;
        acc[at] = gx;
        acc[G + at] = gy;
        acc[2 * G + at] = gz;
        if (acc_abs) acc_abs[at] = sqrt(gx * gx + gy * gy + gz * gz);
      }
}

/* apply_acceleration_to_particles / apply_cell_to_particles
 * p2mesh-kernel.cuh:226-252: value of the particle's cell. */
void orc_mesh_to_particles(int n, long long G, const int* cell_id, const double* cell_acc,
                           const double* cell_phi, double* acc, double* phi) {
  for (int i = 0; i < n; ++i) {
    const long long c = cell_id[i];
    for (int k = 0; k < 3; ++k) acc[(size_t)k * n + i] = cell_acc[k * G + c];
    phi[i] = cell_phi[c];
  }
}
