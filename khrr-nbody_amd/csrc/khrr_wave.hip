// khrr_wave.hip -- damped-wave Poisson stencil, FP32 fields (gfx950).
//
// Replaces wave_diss_iteration (nbody/include/wave-kernel.cuh:5-103) and
// the iteration loop with its triple-buffer swap (nbody/src/nbody.cu:227-239).
// For interior cells
//   phi' = [K*sum6 + 2(1-3K) phi - (1-Q) phi_old - c^2 dt^2 * 4 pi G rho] / (1+Q)
// with K = c^2 dt^2/dx^2, Q = diss*c*dt/2 and faces phi' = 0 (evaluated in
// the algebraically equal increment form, see wave_step_kernel).  The
// dissipation is separable: diss = diss_base + gx(ix) + gy(iy) + gz(iz),
// including the reference's origin quirk (x = ix*dx from 0, so the x term
// always takes its first branch, wave-kernel.cuh:32-46).  The three 1-D
// tables are built on the device in FP64 and rounded to FP32.
//
// Layout: x fastest (AT(x,y,z) = x + y*NX + z*NX^2, common.cuh:77).  A
// 64x4 thread tile marches along z keeping phi(z-1), phi(z), phi(z+1) in
// registers, so each cell reads phi, phi_old and the source once from HBM
// and writes phi' once: 16 B/cell/iteration.  The x/y neighbours are
// re-reads of lines the tile already pulled (L1/L2 hits).
#include <cmath>

#include "khrr_internal.h"

namespace {

constexpr double kPi = 3.14159265358979;  // common.cuh:7
constexpr int kTX = 64, kTY = 4, kZC = 32;

struct WaveConst {
  float half_cdt;  // c*dt/2
  float K;         // c^2 dt^2 / dx^2
  float KF;        // -c^2 dt^2
};

// Per-axis dissipation terms, wave-kernel.cuh:40-68 (FP64, then FP32).
// tab[0..nx) = diss_base + gx, tab[nx..2nx) = gy, tab[2nx..3nx) = gz.
__global__ void wave_tables(int nx, double dx, double diss_base, double diss_extra, double sim_l,
                            double bc_l, float* __restrict__ tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  const double v = i * dx;
  const double edge = sim_l - bc_l;
  double gx = 0.0, gy = 0.0;
  if (v > -edge)
    gx = diss_extra * fabs(v - edge) * fabs(v - edge);
  else if (v < -edge)
    gx = diss_extra * fabs(v + edge) * fabs(v + edge);
  if (v > edge)
    gy = diss_extra * fabs(v - edge) * fabs(v - edge);
  else if (v < -edge)
    gy = diss_extra * fabs(v + edge) * fabs(v + edge);
  tab[i] = (float)(diss_base + gx);
  tab[nx + i] = (float)gy;
  tab[2 * nx + i] = (float)gy;  // z uses the same branch structure as y
}

__global__ __launch_bounds__(256) void wave_source_kernel(size_t G, double scale,
                                                          const double* __restrict__ mass,
                                                          float* __restrict__ src) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < G) src[i] = (float)(scale * mass[i]);
}

__global__ __launch_bounds__(kTX* kTY) void wave_step_kernel(int nx, WaveConst k,
                                                             const float* __restrict__ tab,
                                                             const float* __restrict__ phi,
                                                             const float* __restrict__ phi_old,
                                                             const float* __restrict__ src,
                                                             float* __restrict__ phi_new) {
  const int ix = blockIdx.x * kTX + threadIdx.x;
  const int iy = blockIdx.y * kTY + threadIdx.y;
  if (ix >= nx || iy >= nx) return;
  const int z0 = blockIdx.z * kZC;
  const int z1 = min(nx, z0 + kZC);
  const size_t plane = (size_t)nx * nx;
  const bool xy_face = ix == 0 || ix == nx - 1 || iy == 0 || iy == nx - 1;
  const float dxy = tab[ix] + tab[nx + iy];
  size_t at = (size_t)ix + (size_t)iy * nx + (size_t)z0 * plane;
  float cm = z0 > 0 ? phi[at - plane] : 0.f;
  float cc = phi[at];
  for (int z = z0; z < z1; ++z, at += plane) {
    const float cp = z + 1 < nx ? phi[at + plane] : 0.f;
    float out = 0.f;
    if (!xy_face && z != 0 && z != nx - 1) {
      const float s6 = ((phi[at - 1] + phi[at + 1]) + (phi[at - nx] + phi[at + nx])) + (cm + cp);
      const float diss = dxy + tab[2 * nx + z];
      const float Q = diss * k.half_cdt;
      const float invW = __builtin_amdgcn_rcpf(1.f + Q);
      // Increment form of the same update:
      //   phi' = phi + [K (sum6 - 6 phi) + (1 - Q)(phi - phi_old) + KF src] / (1 + Q)
      // Near steady state the bracket is small, so FP32 rounding applies to
      // the increment instead of to O(phi) terms (20-35x less drift than
      // evaluating the reference's expression term by term in FP32).
      const float lap = __builtin_fmaf(-6.f, cc, s6);
      out = cc + invW * (k.K * lap + (1.f - Q) * (cc - phi_old[at]) + k.KF * src[at]);
    }
    phi_new[at] = out;
    cm = cc;
    cc = cp;
  }
}

int check_wave(khrr_ctx* ctx, const khrr_wave* w) {
  if (!w) return khrr::fail(ctx, KHRR_EINVAL, "wave parameters are NULL");
  if (w->nx < 1 || (long long)w->nx * w->nx * w->nx > (1ll << 31))
    return khrr::fail(ctx, KHRR_EINVAL, "invalid nx %d", w->nx);
  if (!(w->dx > 0.0)) return khrr::fail(ctx, KHRR_EINVAL, "dx must be > 0");
  return 0;
}

WaveConst make_const(const khrr_wave* w) {
  const double cdt = w->c * w->dt;
  const double K = (w->c * w->c) * (w->dt * w->dt) / (w->dx * w->dx);
  WaveConst k;
  k.half_cdt = (float)(0.5 * cdt);
  k.K = (float)K;
  k.KF = (float)(-(w->c * w->c) * (w->dt * w->dt));
  return k;
}

int launch_step(khrr_ctx* ctx, const khrr_wave* w, const WaveConst& k, const float* tab,
                float* phi_new, const float* phi, const float* phi_old, const float* src) {
  const dim3 block(kTX, kTY);
  const dim3 grid(khrr::ceil_div(w->nx, kTX), khrr::ceil_div(w->nx, kTY),
                  khrr::ceil_div(w->nx, kZC));
  hipLaunchKernelGGL(wave_step_kernel, grid, block, 0, ctx->stream, w->nx, k, tab, phi, phi_old,
                     src, phi_new);
  return 0;
}

int build_tables(khrr_ctx* ctx, const khrr_wave* w, float** tab) {
  void* base = nullptr;
  KHRR_TRY(khrr::scratch(ctx, khrr::Carver::need(3 * (size_t)w->nx, sizeof(float)), &base));
  *tab = static_cast<float*>(base);
  hipLaunchKernelGGL(wave_tables, dim3(khrr::ceil_div(w->nx, 256)), dim3(256), 0, ctx->stream,
                     w->nx, w->dx, w->diss_base, w->diss_extra, w->sim_l, w->bc_l, *tab);
  return khrr::check_launch(ctx, "wave_tables");
}

}  // namespace

extern "C" {

int khrr_wave_source(khrr_ctx* ctx, const khrr_wave* w, const double* cell_mass, float* src) {
  KHRR_TRY(khrr::enter(ctx));
  KHRR_TRY(check_wave(ctx, w));
  if (!cell_mass || !src) return khrr::fail(ctx, KHRR_EINVAL, "null field pointer");
  const size_t G = (size_t)w->nx * w->nx * w->nx;
  // f = 4*PI*G*rho, rho = m / dx^3, G = 1 (wave-kernel.cuh:36-38)
  const double scale = 4 * kPi * 1.0 / (w->dx * w->dx * w->dx);
  hipLaunchKernelGGL(wave_source_kernel, dim3(khrr::ceil_div((long long)G, 256)), dim3(256), 0,
                     ctx->stream, G, scale, cell_mass, src);
  return khrr::check_launch(ctx, "wave_source_kernel");
}

int khrr_wave_step(khrr_ctx* ctx, const khrr_wave* w, float* phi_new, const float* phi,
                   const float* phi_old, const float* src) {
  KHRR_TRY(khrr::enter(ctx));
  KHRR_TRY(check_wave(ctx, w));
  if (!phi_new || !phi || !phi_old || !src)
    return khrr::fail(ctx, KHRR_EINVAL, "null field pointer");
  float* tab = nullptr;
  KHRR_TRY(build_tables(ctx, w, &tab));
  const WaveConst k = make_const(w);
  khrr::Timed tm(ctx, KHRR_K_WAVE);
  launch_step(ctx, w, k, tab, phi_new, phi, phi_old, src);
  return khrr::check_launch(ctx, "wave_step_kernel");
}

int khrr_wave_iterate(khrr_ctx* ctx, const khrr_wave* w, float** prev, float** curr,
                      float** next, const float* src, int iters) {
  KHRR_TRY(khrr::enter(ctx));
  KHRR_TRY(check_wave(ctx, w));
  if (!prev || !curr || !next || !*prev || !*curr || !*next || !src)
    return khrr::fail(ctx, KHRR_EINVAL, "null field pointer");
  if (iters < 0) return khrr::fail(ctx, KHRR_EINVAL, "iters < 0");
  float* tab = nullptr;
  KHRR_TRY(build_tables(ctx, w, &tab));
  const WaveConst k = make_const(w);
  float *p = *prev, *c = *curr, *nx = *next;
  for (int it = 0; it < iters; ++it) {
    {
      khrr::Timed tm(ctx, KHRR_K_WAVE);
      launch_step(ctx, w, k, tab, nx, c, p, src);
    }
    KHRR_TRY(khrr::check_launch(ctx, "wave_step_kernel"));
    float* t = p;  // swap(prev, curr); swap(curr, next)
    p = c;
    c = nx;
    nx = t;
  }
  *prev = p;
  *curr = c;
  *next = nx;
  return 0;
}

}  // extern "C"
