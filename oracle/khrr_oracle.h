/*
 * khrr_oracle.h -- CPU FP64 restatement of the KHRR-NBody hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path
 * and the timed CPU baseline; it is never linked into, loaded by or called
 * from the product library (khrr-nbody_amd/).  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py use it.
 *
 * Every function restates one piece of the reference (file:line below) in
 * IEEE FP64 with no FMA contraction (-ffp-contract=off), in the reference's
 * own evaluation order.  Inputs that the GPU path holds in FP32 are passed
 * here widened to FP64, so only arithmetic error is measured.
 *
 * Layout conventions match include/khrr.h: "vec3 planes" are 3*n values
 * (x block, y block, z block); softening is h = eps^2/2.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   orc_accel/orc_phi  two-body closed form, symmetric-polygon zero force,
 *                      momentum balance sum m_i a_i = 0, far field -> M/r^2,
 *                      self term m/sqrt(2h), independent numpy longdouble
 *                      brute force, golden vectors in tests/golden/
 *                      (40-digit mpmath evaluation).
 *   orc_p2mesh_cell_phi numpy per-cell mean of phi, index-order sum.
 *   orc_integrate      constant force integrated exactly, free motion,
 *                      Kepler circular orbit energy, dev_fex continuity at
 *                      Rh2 and Rb and closed forms outside them.
 *   orc_p2mesh_bin     numpy brute-force recount, np.cumsum, block-granular
 *                      restatement, permutation property, mass conservation,
 *                      saturation/NaN special cases.
 *   orc_wave_*         zero-in zero-out, linearity in (phi, phi_old, src),
 *                      faces = 0, independent numpy restatement, single-cell
 *                      impulse response by hand.
 *   orc_diag           closed forms on hand-made states, numpy sums.
 *   orc_mesh_acc       linear fields (exact gradient everywhere), quadratic
 *                      fields (exact inside, closed-form face error),
 *                      numpy.gradient with first-order edges.
 *   orc_mesh_to_particles  numpy fancy indexing.
 *   orc_p2mesh_cell_weighted  constant value -> that value, single-body
 *                      cells exact, conservation sum_c M_c v_c = sum_i m_i v_i,
 *                      linearity, permutation of bodies inside a cell.
 *   orc_p2mesh_cell_avg  equal masses -> the body mass exactly, -1 for an
 *                      empty cell.
 *   orc_acc_abs        Pythagorean quadruples exact, sign and axis
 *                      invariance, numpy.linalg.norm.
 */
#ifndef KHRR_ORACLE_H
#define KHRR_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double Mh, Mh_inf, a, Rh, Mb, b, Rb, Rh2, con, const1, c_phi_h, c_phi_b;
} orc_extpot;

/* nbody.cu:1111-1120 */
void orc_extpot_init(double Mh, double a, double Rh, double Mb, double b, double Rb,
                     orc_extpot* e);
/* nbody-kernel.cuh:8-40 */
void orc_dev_fex(const orc_extpot* e, const double p[3], double f[3]);

/* ACCEL_kernel nbody-kernel.cuh:89-142 and PHI_kernel :43-86 (acc or phi
 * may be NULL).  accumulate != 0: acc += sum, phi -= sum; else overwrite. */
void orc_accel(int n_i, const double* pos_i, const double* h_i, int n_j, const double* pos_j,
               const double* m_j, const double* h_j, double* acc, double* phi, int accumulate);
/* Same, evaluated only for the listed i rows (sample-based force check):
 * row r uses i = rows[r] of pos_i/h_i; outputs are indexed by r. */
void orc_accel_rows(int n_rows, const int* rows, int n_i, const double* pos_i, const double* h_i,
                    int n_j, const double* pos_j, const double* m_j, const double* h_j,
                    double* acc, double* phi);

/* kernelNbody_integTime nbody-kernel.cuh:144-194 (ext may be NULL). */
void orc_integrate(int n, int istep, double dt, double t, const orc_extpot* ext,
                   const double* acc, double* pos_t, double* vel_t, const double* pos,
                   const double* vel, double* acc0);

/* result() sums nbody.cu:797-833; out has 17 entries in the order of
 * KHRR_DIAG_* in include/khrr.h.  phi may be NULL. */
void orc_diag(int n, const double* pos, const double* vel, const double* mass, const double* phi,
              double* out);

/* Reference loop nbody.cu:1468-1594 for `steps` steps of length dt on a
 * state (pos, vel: vec3 planes, updated in place).  compat_accumulate != 0
 * reproduces the never-reset force accumulator of nbody.cu:1560.
 * diag_out (may be NULL) receives 17 diagnostics at t=0 and after every
 * diag_every-th step: (steps/diag_every+1)*17 doubles; phi for EP is
 * recomputed at those points. */
void orc_run(int n, double* pos, double* vel, const double* mass, const double* h, double dt,
             int steps, const orc_extpot* ext, int compat_accumulate, double* diag_out,
             int diag_every);

/* p2mesh binning p2mesh-kernel.cuh:11-137 with the deterministic choices
 * of include/khrr.h (index-order id_in_cell, per-cell and block-granular
 * offsets, index-order FP64 cell mass).  Any output except count may be
 * NULL.  Returns 0, or -1 when the grid is invalid. */
int orc_p2mesh_bin(const double gmin[3], const double gcell[3], const int gsize[3], int n,
                   const double* pos, const double* mass, int* cell_id, int* id_in_cell,
                   int* count, int* offset, int* start_id_ref, int* block_incl,
                   double* cell_mass);
/* computeCellPhiUnsorted p2mesh-kernel.cuh:152-167 in index order. */
void orc_p2mesh_cell_phi(int n, long long G, const int* cell_id, const int* count,
                         const double* phi, double* cell_phi, int accumulate);
/* apply_particle_to_cell_weighted p2mesh-kernel.cuh:273-287 in index
 * order from 0: cell_value[c] += value[i] * (mass[i] / cell_mass[c]). */
void orc_p2mesh_cell_weighted(int n, long long G, const int* cell_id, const double* cell_mass,
                              const double* mass, const double* value, double* cell_value);
/* apply_cell_to_particles_avg p2mesh-kernel.cuh:256-271: cell_value/count,
 * -1 for a cell of count 0. */
void orc_p2mesh_cell_avg(int n, const int* cell_id, const int* count, const double* cell_value,
                         double* out);
/* acc_abs p2mesh-kernel.cuh:289-298 of vec3 planes (norm3, common.cuh:84-96). */
void orc_acc_abs(int n, const double* v, double* out);
/* (int) conversion with GPU cvt semantics: truncate, saturate, NaN -> 0. */
int orc_sat_trunc(double v);

/* wave_diss_iteration wave-kernel.cuh:5-103 in FP64 (mass is cell mass). */
void orc_wave_step(int nx, double dx, double c, double dt, double diss_base, double diss_extra,
                   double sim_l, double bc_l, double* phi_new, const double* phi,
                   const double* phi_old, const double* mass);
/* `iters` steps with the rotation of nbody.cu:227-239; returns in phi_out
 * the final "curr" buffer (and the final "prev" in prev_out if non-NULL).
 * prev/curr are the initial fields (copied, not modified). */
void orc_wave_iterate(int nx, double dx, double c, double dt, double diss_base,
                      double diss_extra, double sim_l, double bc_l, const double* prev0,
                      const double* curr0, const double* mass, int iters, double* phi_out,
                      double* prev_out);

/* acc_field p2mesh-kernel.cuh:169-224 on the full 3-D grid (FP64). */
void orc_mesh_acc(int nx, double dx, const double* phi, double* acc, double* acc_abs);
/* apply_acceleration_to_particles / apply_cell_to_particles :226-252. */
void orc_mesh_to_particles(int n, long long G, const int* cell_id, const double* cell_acc,
                           const double* cell_phi, double* acc, double* phi);

/* Threads OpenMP will use (for the baseline's core count). */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
