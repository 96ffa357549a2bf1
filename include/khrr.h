/*
 * khrr.h -- C-ABI of the MI355X (gfx950) N-body hot path.
 *
 * Drop-in replacement for the device kernels of KHRR-NBody that the host
 * driver launches through RR's CuCall(kernel, grid, block)(args...):
 *
 *   ACCEL_kernel            nbody/include/nbody-kernel.cuh:89-142   -> khrr_accel
 *   PHI_kernel              nbody/include/nbody-kernel.cuh:43-86    -> khrr_accel (acc == NULL)
 *   kernelNbody_integTime   nbody/include/nbody-kernel.cuh:144-194  -> khrr_integrate
 *   dev_fex + DataBlock     nbody/include/nbody-kernel.cuh:8-40,
 *                           nbody/src/nbody.cu:1111-1120            -> khrr_extpot_init
 *   assignParticlesToCells  nbody/include/p2mesh-kernel.cuh:11-57  \
 *   computePrefixSums       nbody/include/p2mesh-kernel.cuh:62-105  |-> khrr_p2mesh_bin
 *   adjustGlobalPrefixSums  nbody/include/p2mesh-kernel.cuh:110-123 |   (the sequence of
 *   computeCellMassesUnsorted p2mesh-kernel.cuh:126-137            /    nbody.cu:139-181)
 *   wave_diss_iteration     nbody/include/wave-kernel.cuh:5-103     -> khrr_wave_step
 *   iteration loop + swap   nbody/src/nbody.cu:227-239              -> khrr_wave_iterate
 *   result() host sums      nbody/src/nbody.cu:797-833              -> khrr_diag
 *
 * Conventions
 *   - Every entry point returns 0 on success or a negative KHRR_E* code; the
 *     message is available from khrr_last_error(ctx).  Nothing calls exit().
 *   - All array arguments are DEVICE pointers owned by the caller, unless
 *     the comment says "host".  Work is enqueued on the context's stream and
 *     is asynchronous unless the comment says "synchronous".
 *   - "vec3 planes": a 3-component per-particle quantity of n entries is one
 *     buffer of 3*n elements, x in [0,n), y in [n,2n), z in [2n,3n).
 *     (The reference stores double3 AoS; planes give coalesced SoA access.)
 *   - Softening is passed as h = eps^2/2 per particle, so the reference's
 *     per-pair 0.5*(eps2_i + eps2_j) (nbody-kernel.cuh:78,126) is h_i + h_j.
 *   - Gravity arithmetic is FP32; the integrator state, binning index math,
 *     cell masses and diagnostics are FP64 (the reference's `real`).
 *   - A context is not re-entrant: calls on one context must not overlap
 *     from several host threads.  Distinct contexts are independent.
 */
#ifndef KHRR_H
#define KHRR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KHRR_ABI_VERSION 2

/* ---- status codes ---------------------------------------------------- */
#define KHRR_OK 0
#define KHRR_EINVAL (-1)   /* bad argument (size, null pointer, flags)      */
#define KHRR_EDEVICE (-2)  /* HIP runtime error                             */
#define KHRR_ENOMEM (-3)   /* device allocation failed / budget exceeded    */
#define KHRR_ESTATE (-4)   /* call not valid in the object's current state  */

/* ---- flags for khrr_accel -------------------------------------------- */
/* acc += sum, phi -= sum (read-modify-write exactly as ACCEL_kernel
 * nbody-kernel.cuh:137-141 and PHI_kernel :85).  Without it the outputs are
 * overwritten: acc = sum, phi = -sum. */
#define KHRR_ACCUMULATE 1u
/* Compensated (TwoSum) accumulation of the per-particle sums: tile partial
 * sums are folded into a (hi, lo) pair so the FP32 result carries the
 * rounding of the pair terms only, not of an N-term running sum. */
#define KHRR_COMPENSATED 2u
/* khrr_accel_batched only: the call's launch sequence is captured into a
 * HIP graph, cached per (composition, pointers, flags, launch shape), and
 * later identical calls replay it with one hipGraphLaunch.  The first call
 * of a key runs eagerly and captures on a private stream; results are the
 * eager results bit for bit.  Up to 4 graphs are kept per context (least
 * recently used evicted). */
#define KHRR_GRAPH 4u

typedef struct khrr_ctx khrr_ctx;

/* ---- context --------------------------------------------------------- */
int khrr_abi_version(void);
/* Creates a context on HIP device `device`; work goes to the null stream
 * until khrr_ctx_set_stream is called. */
int khrr_ctx_create(int device, khrr_ctx** out);
void khrr_ctx_destroy(khrr_ctx* ctx);
/* stream: a hipStream_t (NULL = the null stream). */
int khrr_ctx_set_stream(khrr_ctx* ctx, void* stream);
/* Last error message of this context (or of the calling thread when ctx is
 * NULL).  Valid until the next call on the same context. */
const char* khrr_last_error(const khrr_ctx* ctx);
/* Blocks until all work enqueued by this context has finished. */
int khrr_ctx_sync(khrr_ctx* ctx);
/* Kernel timing.  When enabled, every launch of a timed kernel class is
 * bracketed by HIP events on the context's stream.  khrr_ctx_kernel_time
 * is synchronous and returns the launch count and summed device time (ms)
 * since the last reset. */
#define KHRR_K_GRAVITY 0
#define KHRR_K_INTEGRATE 1
#define KHRR_K_BIN 2
#define KHRR_K_WAVE 3
#define KHRR_K_DIAG 4
#define KHRR_K_COUNT 5
int khrr_ctx_set_timing(khrr_ctx* ctx, int enable);
int khrr_ctx_kernel_time(khrr_ctx* ctx, int kernel_class, long* launches, double* total_ms);
int khrr_ctx_reset_timing(khrr_ctx* ctx);

/* ---- gravity (a1, a2) ------------------------------------------------ */
/* Tunable launch shape of the direct-sum kernel.  Zeroed fields mean
 * "pick automatically" (see khrr_tune_gravity). */
typedef struct {
  int i_per_lane;   /* i-particles held per lane: 1, 2, 4 or 8          */
  int waves_per_block; /* 1..4                                          */
  int j_split;      /* number of independent j-ranges (partial sums)    */
  int j_tile;       /* j count per FP32 partial-sum tile (multiple of 16) */
  int symmetric;    /* 1: when the i-set and j-set are the same buffers,
                       evaluate each unordered pair once and apply it to
                       both bodies (needs i_per_lane 4 or 8 and a partial
                       buffer of 16 B per body per superblock, at most
                       40 GiB; otherwise the one-sided kernel runs)      */
  int superblock;   /* pair-symmetric bodies per superblock: 1024, 2048
                       or 4096 (1024 always runs 4 i per lane), or 0 for
                       the rule of n: 4096 from n = 524288 on, 2048 from
                       32768, 1024 below.  Added in ABI version 2.       */
} khrr_tune;

/* Softened direct sum.  For every i in [0,n_i):
 *   d_ij = pos_j[j] - pos_i[i],  r2 = |d_ij|^2 + h_i[i] + h_j[j]
 *   acc[i] (+)= sum_j m_j d_ij r2^-3/2        (ACCEL_kernel :89-142)
 *   phi[i] (-)= sum_j m_j r2^-1/2, j = i incl. (PHI_kernel :43-86)
 * pos_i: vec3 planes of n_i floats; pos_j: vec3 planes of n_j floats;
 * acc: vec3 planes of n_i floats or NULL; phi: n_i floats or NULL (at least
 * one of acc/phi non-NULL).  The reference requires N % 512 == 0 and reads
 * out of bounds otherwise; here any n_i >= 0, n_j >= 0 is valid.
 * tune may be NULL.  i- and j-sets may alias (same buffers). */
int khrr_accel(khrr_ctx* ctx, int n_i, const float* pos_i, const float* h_i, int n_j,
               const float* pos_j, const float* m_j, const float* h_j, float* acc, float* phi,
               unsigned flags, const khrr_tune* tune);

/* Batched form: n_sys independent systems stored back to back.  System s
 * owns particles [off[s], off[s+1]) of the n_total = off[n_sys] bodies and
 * interacts only with itself.  host_sys_off is the HOST array of the n_sys+1
 * prefix offsets; sys_off, a device copy of it, is accepted for interface
 * symmetry and may be NULL (the work list is built on the host).  pos/acc
 * are vec3 planes of n_total entries.  Per-system rules, functions of the
 * system's own n only:
 *  - n >= 16384 (when its partial slots fit, as for khrr_accel): the
 *    pair-symmetric path, one workgroup per (A, B) superblock item; the
 *    result is bit-identical to khrr_accel with i_per_lane 8, symmetric 1
 *    (superblocks of 1024 bodies below 32768, which run 4 i per lane);
 *  - otherwise the j-range is cut into k = min(32, ceil(n / 1024)) pieces
 *    and (i-group, j-piece) items run longest system first in one
 *    persistent launch, pieces summed in piece order; the result is
 *    bit-identical to khrr_accel with j_split k, j_tile 256, symmetric 0.
 * So a system's result never depends on what else is in the batch.  When
 * both kinds are present the one-sided launch runs on a second stream of
 * the context, concurrently with the pair-symmetric items, and the call's
 * stream waits for it before returning control of its outputs.
 * tune == NULL: the shape khrr_tune_batched stored in the context, else
 * i_per_lane the widest of 4, 2, 1 that gives at least two one-sided
 * items per SIMD (results do not depend on i_per_lane). */
int khrr_accel_batched(khrr_ctx* ctx, int n_sys, const int* sys_off, const int* host_sys_off,
                       const float* pos, const float* h, const float* m, float* acc, float* phi,
                       unsigned flags, const khrr_tune* tune);

/* Project-written autotuner: times candidate launch shapes of khrr_accel
 * on synthetic bodies of size n on this context's device (synchronous) and
 * returns the fastest: one-sided i_per_lane x waves_per_block x j_split
 * with 256-body j-tiles, then j_tile 128/512/1024 for the fastest of
 * those; pair-symmetric i_per_lane 4/8 x superblock 1024/2048/4096,
 * compared per pair-interaction.  Results are cached per power-of-two size
 * bucket in the context; khrr_accel with tune == NULL uses the cache (or a
 * static heuristic when the bucket was never tuned). */
int khrr_tune_gravity(khrr_ctx* ctx, int n, khrr_tune* best, double* best_ms);
/* Same for the batched form on the composition host_sys_off[0..n_sys]
 * (i_per_lane, waves_per_block, then j_tile; j_split and the superblock
 * follow the per-system rules there).  The result is what
 * khrr_accel_batched uses when its tune argument is NULL. */
int khrr_tune_batched(khrr_ctx* ctx, int n_sys, const int* host_sys_off, khrr_tune* best,
                      double* best_ms);
/* Static choice the library makes for n when nothing is tuned. */
int khrr_tune_default(khrr_ctx* ctx, int n, khrr_tune* out);
/* The shape khrr_accel uses for n bodies when its tune argument is NULL,
 * if one was tuned or loaded for n's size bucket (ceil(log2 n));
 * KHRR_ESTATE otherwise.  khrr_tune_batched_cached: the same for
 * khrr_accel_batched. */
int khrr_tune_cached(khrr_ctx* ctx, int n, khrr_tune* out);
int khrr_tune_batched_cached(khrr_ctx* ctx, khrr_tune* out);
/* The autotuner's cache on disk.  khrr_tune_save writes every tuned shape
 * of the context (per size bucket, with its trial time, and the batched
 * shape) as JSON, tagged with the device's architecture name and CU count.
 * khrr_tune_load reads such a file into the context: KHRR_ESTATE when it
 * was made on another kind of device, KHRR_EINVAL when it is malformed or
 * holds an invalid shape (then nothing is loaded).  Host-side, no device
 * work. */
int khrr_tune_save(khrr_ctx* ctx, const char* path);
int khrr_tune_load(khrr_ctx* ctx, const char* path);

/* ---- integrator (a3) ------------------------------------------------- */
/* DataBlock fields used by dev_fex (common.cuh:40-56) */
typedef struct {
  double Mh, Mh_inf, a, Rh, Mb, b, Rb, Rh2, con, const1, c_phi_h, c_phi_b;
} khrr_extpot;
/* Host-side constants exactly as nbody.cu:1111-1120 (Mh = Mb = 0 gives an
 * identically zero external force). */
int khrr_extpot_init(double Mh, double a, double Rh, double Mb, double b, double Rb,
                     khrr_extpot* out);
/* Host evaluation of dev_fex(p) (nbody-kernel.cuh:8-40), for tests. */
int khrr_extpot_force(const khrr_extpot* ext, const double p[3], double f[3]);

/* kernelNbody_integTime (nbody-kernel.cuh:144-194), FP64 state, f = acc +
 * dev_fex(pos):
 *   istep 0 (predictor): vel_t = vel + dt f;  pos_t = pos + dt/2 (vel + vel_t)
 *   istep 1 (corrector): vel_t = (vel_t + vel + dt f)/2;  pos_t = pos;
 *                        acc0 = acc
 * pos_t, vel_t, pos, vel: vec3 planes of n doubles; acc, acc0: vec3 planes
 * of n floats.  pos32 (optional, vec3 planes of n floats) receives
 * (float)pos_t for the next gravity call.  ext may be NULL (no external
 * potential).  t is accepted for signature parity; dev_fex ignores it. */
int khrr_integrate(khrr_ctx* ctx, int n, int istep, double dt, double t, const khrr_extpot* ext,
                   const float* acc, double* pos_t, double* vel_t, const double* pos,
                   const double* vel, float* acc0, float* pos32);

/* ---- diagnostics (a9) ------------------------------------------------ */
#define KHRR_DIAG_EK 0     /* sum m v^2        (nbody.cu:824)               */
#define KHRR_DIAG_EP 1     /* sum m phi        (nbody.cu:825)               */
#define KHRR_DIAG_PX 2     /* momentum         (nbody.cu:821-823)           */
#define KHRR_DIAG_PY 3
#define KHRR_DIAG_PZ 4
#define KHRR_DIAG_LX 5     /* angular momentum (nbody.cu:818-820)           */
#define KHRR_DIAG_LY 6
#define KHRR_DIAG_LZ 7
#define KHRR_DIAG_M 8      /* total mass                                    */
#define KHRR_DIAG_MX 9     /* sum m x (centre of mass * M)                  */
#define KHRR_DIAG_MY 10
#define KHRR_DIAG_MZ 11
#define KHRR_DIAG_VR_MAX 12  /* max |v_r|  (nbody.cu:827-831)               */
#define KHRR_DIAG_VFI_MAX 13 /* max v_phi (signed, starts at 0)             */
#define KHRR_DIAG_VZ_MAX 14
#define KHRR_DIAG_R_MAX 15
#define KHRR_DIAG_Z_MAX 16
#define KHRR_DIAG_COUNT 17
/* FP64 reductions on the device over n particles (deterministic fixed-order
 * tree).  pos, vel: vec3 planes of n doubles; mass: n doubles; phi: n
 * floats (may be NULL: EP = 0).  out: HOST array of KHRR_DIAG_COUNT
 * doubles.  Synchronous. */
int khrr_diag(khrr_ctx* ctx, int n, const double* pos, const double* vel, const double* mass,
              const float* phi, double* out);

/* ---- p2mesh binning (a5-a7) ------------------------------------------ */
typedef struct {
  double min[3];    /* domainMin  (WaveEqData, common.cuh:58)            */
  double cell[3];   /* cellSize                                           */
  int size[3];      /* gridSize;  cells = size[0]*size[1]*size[2]         */
} khrr_grid;
#define KHRR_BIN_BLOCK 512 /* BLOCK_SIZE of the reference's cell blocks   */
/* Bins n particles (pos: vec3 planes of n doubles, mass: n doubles).
 *   cell_id[i]     = iz*sx*sy + iy*sz + ix with i* = clamp((int)((p-min)/cell),
 *                    0, size-1) in FP64, saturating conversion, NaN -> 0, z
 *                    clamped with size[1]   (p2mesh-kernel.cuh:25-43)
 *   id_in_cell[i]  = rank of i among the particles of its cell in index order
 *                    (the reference's atomic order is unspecified, :49)
 *   count[c]       = particles in cell c                      (:55)
 *   offset[c]      = exclusive prefix sum of count (per cell)
 *   start_id_ref[c]= sum of counts of all 512-cell blocks before block(c),
 *                    the value the reference's CellInfo.start_id holds after
 *                    computePrefixSums + adjustGlobalPrefixSums (:62-123)
 *   block_incl[b]  = inclusive prefix of 512-cell block totals (maxPBC)
 *   cell_mass[c]   = FP64 sum of mass over the cell in index order (:126-137)
 * Any output except count may be NULL.  Scratch is owned by the context. */
int khrr_p2mesh_bin(khrr_ctx* ctx, const khrr_grid* grid, int n, const double* pos,
                    const double* mass, int* cell_id, int* id_in_cell, int* count, int* offset,
                    int* start_id_ref, int* block_incl, double* cell_mass);

/* Allocation-free form of khrr_p2mesh_bin: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops): same arguments, kernels and
 * results as khrr_p2mesh_bin, but the scratch is one fixed device array of
 * the library (about 21 MB per device, held from the first context on that
 * device until the process ends) instead of the context's growing arena,
 * and no timers are recorded.  Limits: n <= 1048576 bodies and at most
 * 16777216 (256^3) cells; larger calls fail with KHRR_EINVAL.  The array
 * is shared by every call on one device: calls that overlap on one device
 * race and must be serialised by the caller.  Calls no hipMalloc/hipFree. */
int khrr_p2mesh_bin_noalloc(khrr_ctx* ctx, const khrr_grid* grid, int n, const double* pos,
                            const double* mass, int* cell_id, int* id_in_cell, int* count,
                            int* offset, int* start_id_ref, int* block_incl, double* cell_mass);

/* computeCellPhiUnsorted (p2mesh-kernel.cuh:152-167): for every particle i
 * of cell c, cell_phi[c] += phi[i] / count[c], summed in FP64 in particle
 * index order (the reference's FP64 atomics have no fixed order).  Uses
 * the cell_id, id_in_cell, count and offset arrays of khrr_p2mesh_bin.
 * accumulate == 0 starts every cell from 0; accumulate != 0 adds to the
 * existing cell_phi, which is what the reference's save path does with a
 * buffer it never clears (nbody.cu:1628-1634).  cell_phi: G doubles. */
int khrr_p2mesh_cell_phi(khrr_ctx* ctx, int n, long long n_cells, const int* cell_id,
                         const int* id_in_cell, const int* count, const int* offset,
                         const float* phi, double* cell_phi, int accumulate);
/* apply_particle_to_cell_weighted (p2mesh-kernel.cuh:273-287): for every
 * particle i of cell c, cell_value[c] += value[i] * (mass[i] / cell_mass[c]),
 * summed in FP64 in particle index order from 0 (the reference adds with
 * FP64 atomics into a fresh buffer; nbody.cu:545,585).  cell_mass is the
 * binning's FP64 cell mass; cells of no bodies get 0.  Same binning inputs
 * as khrr_p2mesh_cell_phi.  value, mass: n doubles; cell_value: G doubles. */
int khrr_p2mesh_cell_weighted(khrr_ctx* ctx, int n, long long n_cells, const int* cell_id,
                              const int* id_in_cell, const int* count, const int* offset,
                              const double* cell_mass, const double* mass, const double* value,
                              double* cell_value);
/* apply_cell_to_particles_avg (p2mesh-kernel.cuh:256-271): out[i] =
 * cell_value[c] / count[c] for the particle's cell c, -1 where count[c] is
 * 0, NaN where c is outside [0, n_cells).  out: n doubles. */
int khrr_p2mesh_cell_avg(khrr_ctx* ctx, int n, long long n_cells, const int* cell_id,
                         const int* count, const double* cell_value, double* out);

/* ---- damped-wave Poisson stencil (a8) -------------------------------- */
typedef struct {
  int nx;            /* cells per edge (cubic grid, AT() index x fastest) */
  double dx;         /* dx_wave   */
  double c;          /* c_wave    */
  double dt;         /* dt_wave   */
  double diss_base;
  double diss_extra;
  double sim_l;
  double bc_l;
} khrr_wave;
/* src[c] = (float)(4*PI*G*cell_mass[c]/dx^3), G = 1 (wave-kernel.cuh:36-38) */
int khrr_wave_source(khrr_ctx* ctx, const khrr_wave* w, const double* cell_mass, float* src);
/* One wave_diss_iteration: phi_new from phi, phi_old, src (FP32, nx^3). */
int khrr_wave_step(khrr_ctx* ctx, const khrr_wave* w, float* phi_new, const float* phi,
                   const float* phi_old, const float* src);
/* `iters` iterations with the reference's buffer rotation (nbody.cu:227-239):
 *   next = step(curr, prev); prev <- curr; curr <- next; next <- old prev.
 * The three pointers are updated in place (host pointers to device pointers),
 * so afterwards *curr holds the newest field and *prev the one before. */
int khrr_wave_iterate(khrr_ctx* ctx, const khrr_wave* w, float** prev, float** curr,
                      float** next, const float* src, int iters);

/* ---- mesh force of the wave solver ------------------------------------ */
/* acc_field (p2mesh-kernel.cuh:169-224): cell_acc = -grad(cell_phi) of an
 * nx^3 field (x fastest): 0.5 (phi[+1] - phi[-1]) inside, one-sided
 * differences on the faces, times -1/dx; evaluated in FP64 from the FP32
 * field and rounded to FP32.  cell_acc: vec3 planes of nx^3 floats;
 * cell_acc_abs: nx^3 floats (|a|, norm3 of common.cuh:94) or NULL.  The
 * reference launches this kernel over a 1-D grid (nbody.cu:244), which
 * evaluates every cell as if iy = iz = 0 and reads past the field; here
 * every cell gets its own 3-D stencil.  nx >= 2. */
int khrr_mesh_acc(khrr_ctx* ctx, int nx, double dx, const float* cell_phi, float* cell_acc,
                  float* cell_acc_abs);
/* The same field of an FP64 potential, in FP64 throughout: the form the
 * debug dumps of convert_particles apply to the FP64 cell potential of the
 * bodies (nbody.cu:484-517).  cell_acc: vec3 planes of nx^3 doubles;
 * cell_acc_abs: nx^3 doubles or NULL. */
int khrr_mesh_acc_f64(khrr_ctx* ctx, int nx, double dx, const double* cell_phi,
                      double* cell_acc, double* cell_acc_abs);
/* acc_abs (p2mesh-kernel.cuh:289-298): out[i] = |v_i| of FP32 vec3 planes,
 * evaluated in FP64 as sqrt((x*x + y*y) + z*z) (norm3, common.cuh:84-96).
 * v: vec3 planes of n floats; out: n doubles. */
int khrr_acc_abs(khrr_ctx* ctx, int n, const float* v, double* out);
/* apply_acceleration_to_particles + apply_cell_to_particles
 * (p2mesh-kernel.cuh:226-252): acc[.][i] = cell_acc[.][cell_id[i]] and
 * phi[i] = cell_phi[cell_id[i]] (cell_id from khrr_p2mesh_bin; an index
 * outside [0, n_cells) gives NaN).  acc: vec3 planes of n floats or NULL;
 * phi: n floats or NULL (not both). */
int khrr_mesh_to_particles(khrr_ctx* ctx, int n, long long n_cells, const int* cell_id,
                           const float* cell_acc, const float* cell_phi, float* acc, float* phi);

/* ---- serving form: continuous batching ------------------------------- */
/* No counterpart in the reference (one simulation per process).  A server
 * keeps every admitted simulation's state resident in device memory and
 * advances all of them together, one predictor -> gravity -> corrector
 * step per khrr_server_step, with a single batched gravity launch over
 * every active system (khrr_accel_batched's per-system rules; each round
 * picks i_per_lane for its own composition as khrr_accel_batched does
 * with a NULL tune, starting from the tuned one).  Requests
 * that arrive while others run are admitted
 * (first come, first served) as soon as device memory within the budget
 * and an active slot are free.  A finished simulation returns its
 * conservation-law diagnostics at t = 0 and at the end, and the FP64 mass
 * of its bodies binned on the requested grid.  Every system's arithmetic
 * depends only on its own bodies, so a result is bit-identical whether the
 * request ran alone or among others. */
typedef struct khrr_server khrr_server;
typedef struct {
  int n;                /* bodies                                            */
  const double* pos;    /* HOST vec3 planes of n doubles                     */
  const double* vel;    /* HOST vec3 planes of n doubles                     */
  const double* mass;   /* HOST n doubles                                    */
  const double* eps;    /* HOST n softening lengths (h = eps^2 / 2)          */
  double dt;            /* step                                              */
  int steps;            /* predictor-corrector steps to run (tmax / dt)      */
  khrr_grid grid;       /* grid of the returned density                     */
} khrr_request;
typedef struct {
  int active;           /* simulations on the device                        */
  int pending;          /* submitted, waiting for admission                 */
  int finished;         /* finished, result not yet collected               */
  long long rounds;     /* khrr_server_step calls that ran a step           */
  double pair_interactions; /* N^2 per system per force evaluation, total   */
  size_t bytes_in_use;  /* device bytes held by admitted simulations        */
  size_t bytes_budget;
} khrr_server_info;
/* hbm_budget: device bytes the server may hold for simulation state: a
 * simulation of n bodies holds 152 B per body (FP32 gravity arrays and FP64
 * state) plus the batched gravity's partial sums for it: 16 B per body per
 * superblock when n >= 16384 (1024 bodies below 32768, 2048 below 524288,
 * 4096 from there on), else 16 B per body per j-piece when there are two
 * or more.  The
 * FP32 arena for budget / 152 bodies is reserved at creation.
 * max_active: simultaneous simulations (>= 1). */
int khrr_server_create(khrr_ctx* ctx, size_t hbm_budget, int max_active, khrr_server** out);
/* Frees the server's device memory; call it before khrr_ctx_destroy of its
 * context (the server works on the context's stream). */
void khrr_server_destroy(khrr_server* srv);
/* Round scheduling.  pairs == 0 (the default): every active simulation
 * takes one step per round.  pairs > 0: the active simulations are taken
 * in order of least remaining work (n^2 x steps left, ties by id) until
 * the round holds at least `pairs` pair-interactions per step (always at
 * least one); the others wait for a later round.  Results do not depend on
 * the schedule. */
int khrr_server_set_round_pairs(khrr_server* srv, double pairs);
/* Copies the request's host arrays and queues it; *id identifies it.
 * Rejected (KHRR_ENOMEM) when it can never fit the budget. */
int khrr_server_submit(khrr_server* srv, const khrr_request* req, long long* id);
/* One scheduling round: admit what fits, advance every active simulation
 * by one step, finish those that reached their step count.  Writes up to
 * max_done finished ids to done_ids (host) and their number to *n_done.
 * A round that admits or finishes a simulation returns when the device has
 * completed it (its results are then ready); any other round returns once
 * its launches are queued, with at most two rounds in flight, so the host
 * prepares the next round while the device runs this one. */
int khrr_server_step(khrr_server* srv, long long* done_ids, int max_done, int* n_done);
/* Copies a finished simulation's results and releases it.  diag0/diag1:
 * KHRR_DIAG_COUNT doubles each (t = 0 and final; EP uses the potential);
 * cell_mass: HOST array of grid.size[0]*size[1]*size[2] doubles or NULL. */
int khrr_server_result(khrr_server* srv, long long id, double* diag0, double* diag1,
                       double* cell_mass, size_t cell_mass_len);
int khrr_server_stats(khrr_server* srv, khrr_server_info* out);

#ifdef __cplusplus
}
#endif
#endif /* KHRR_H */
