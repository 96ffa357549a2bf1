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
// 64x16 thread tile marches along z keeping phi(z-1), phi(z), phi(z+1) in
// registers, so each cell reads phi, phi_old and the source once from HBM
// and writes phi' once: 16 B/cell/iteration.  The x/y neighbours are
// re-reads of lines the tile already pulled (L1/L2 hits).
#include <cmath>

#include "khrr_internal.h"

// Every fused multiply-add below is written out, so the one-step and the
// two-step kernels round identically.
#pragma clang fp contract(off)

namespace {

constexpr double kPi = 3.14159265358979;  // common.cuh:7
// One-step kernel: a 64 x 16 column tile marching 16 planes (47.0 us per
// iteration at 256^3, against 55.0 us for 64 x 4 columns marching 32).
constexpr int kTX = 64, kTY = 16, kZC = 16;
// Two-step kernel: output tile kFX x kFY x kFZ; first-step footprint with a
// one-cell ring, input plane with a two-cell ring.  At 256^3 a 64 x 32 tile
// marching 16 planes with 768 threads (66 KB of LDS, two workgroups per
// CU) took 42.0 us per iteration; 64 x 16 with 256 threads 44.1 us, and
// other heights, runs and widths 43-68 us.
#ifndef KHRR_WAVE_TILE_Y
#define KHRR_WAVE_TILE_Y 32
#endif
#ifndef KHRR_WAVE_TILE_Z
#define KHRR_WAVE_TILE_Z 16
#endif
constexpr int kFX = 64, kFY = KHRR_WAVE_TILE_Y, kFZ = KHRR_WAVE_TILE_Z, kFT = 768;
constexpr int kQW = kFX + 2, kQH = kFY + 2;
constexpr int kPW = kFX + 4, kPH = kFY + 4;

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

// One interior cell: cc = phi, s6 = ((x- + x+) + (y- + y+)) + (z- + z+),
// in the increment form of the reference's update
//   phi' = phi + [K (s6 - 6 phi) + (1 - Q)(phi - phi_old) + KF src] / (1 + Q),
// Q = diss * c dt / 2.  Near steady state the bracket is small, so FP32
// rounding applies to the increment instead of to O(phi) terms (20-35x
// less drift than evaluating the reference's expression term by term).
__device__ __forceinline__ float wave_update(float cc, float s6, float old, float src, float diss,
                                             const WaveConst& k) {
  const float Q = diss * k.half_cdt;
  const float invW = __builtin_amdgcn_rcpf(1.f + Q);
  const float lap = __builtin_fmaf(-6.f, cc, s6);
  const float br = __builtin_fmaf(k.K, lap, __builtin_fmaf(1.f - Q, cc - old, k.KF * src));
  return __builtin_fmaf(invW, br, cc);
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
      out = wave_update(cc, s6, phi_old[at], src[at], dxy + tab[2 * nx + z], k);
    }
    phi_new[at] = out;
    cm = cc;
    cc = cp;
  }
}

__device__ __forceinline__ int mod3(int v) { return ((v % 3) + 3) % 3; }

// Two iterations in one pass (temporal blocking): from p0 = phi_old and
// p1 = phi, writes p2 = step(p1, p0) and p3 = step(p2, p1), each cell by the
// same arithmetic as wave_step_kernel.  A workgroup marches along z over
// kFZ output planes; per plane it stages p1 with a two-cell ring in LDS
// (ring of 4 planes), computes the first step on the tile plus a one-cell
// ring into LDS (ring of 3 planes), and the second step on the tile one
// plane behind.  Global loads for the next plane are issued into registers
// before the current plane is computed.  p2 and p3 must not alias p0 or p1
// (neighbouring tiles read their rings).  HBM traffic: read p0, p1, src,
// write p2, p3 = 20 B per cell per two iterations (plus ring re-reads,
// mostly from L2).
constexpr int kLoadPer = (kPW * kPH + kFT - 1) / kFT;
constexpr int kQPer = (kQW * kQH + kFT - 1) / kFT;
constexpr int kOPer = (kFX * kFY + kFT - 1) / kFT;

__global__ __launch_bounds__(kFT) void wave_step2_kernel(int nx, WaveConst k,
                                                         const float* __restrict__ tab,
                                                         const float* __restrict__ p0,
                                                         const float* __restrict__ p1,
                                                         const float* __restrict__ src,
                                                         float* __restrict__ p2,
                                                         float* __restrict__ p3) {
  __shared__ float P[4][kPH * kPW];
  __shared__ float Qs[3][kQH * kQW];
  // Workgroups are dealt to the 8 XCDs round-robin; tile t = (b % 8) * per
  // + b / 8 gives each XCD a contiguous run of tiles (x fastest, then y,
  // then z), so the rings a tile re-reads were mostly pulled into the same
  // XCD's L2 by its neighbours.
  const int nbx = (nx + kFX - 1) / kFX, nby = (nx + kFY - 1) / kFY, nbz = (nx + kFZ - 1) / kFZ;
  const int per = (nbx * nby * nbz + 7) / 8;
  const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (t >= nbx * nby * nbz) return;
  const int x0 = (t % nbx) * kFX, y0 = ((t / nbx) % nby) * kFY;
  const int z0 = (t / (nbx * nby)) * kFZ;
  const int z1 = min(nx, z0 + kFZ);
  const size_t plane = (size_t)nx * nx;
  const int tid = threadIdx.x;

  // Per-thread cells, fixed for the whole march: xy offsets in the field
  // (-1: none), LDS slots and the xy part of the dissipation.
  int lg[kLoadPer];  // staged input cells (tile + two-cell ring)
  int qg[kQPer];     // first-step cells (tile + one-cell ring), interior only
  float qd[kQPer];
  int og[kOPer];     // tile cells
  bool oi[kOPer];    // tile cell is xy-interior
  float od[kOPer];
#pragma unroll
  for (int j = 0; j < kLoadPer; ++j) {
    const int f = tid + j * kFT;
    const int py = f / kPW, px = f - py * kPW;
    const int gx = x0 - 2 + px, gy = y0 - 2 + py;
    lg[j] = (f < kPW * kPH && gx >= 0 && gx < nx && gy >= 0 && gy < nx) ? gy * nx + gx : -1;
  }
#pragma unroll
  for (int j = 0; j < kQPer; ++j) {
    const int f = tid + j * kFT;
    const int qy = f / kQW, qx = f - qy * kQW;
    const int gx = x0 - 1 + qx, gy = y0 - 1 + qy;
    const bool in = f < kQW * kQH && gx >= 1 && gx <= nx - 2 && gy >= 1 && gy <= nx - 2;
    qg[j] = in ? gy * nx + gx : -1;
    qd[j] = in ? tab[gx] + tab[nx + gy] : 0.f;
  }
#pragma unroll
  for (int j = 0; j < kOPer; ++j) {
    const int f = tid + j * kFT;
    const int oy = f / kFX, ox = f - oy * kFX;
    const int gx = x0 + ox, gy = y0 + oy;
    og[j] = (f < kFX * kFY && gx < nx && gy < nx) ? gy * nx + gx : -1;
    oi[j] = gx >= 1 && gx <= nx - 2 && gy >= 1 && gy <= nx - 2;
    od[j] = og[j] >= 0 ? tab[gx] + tab[nx + gy] : 0.f;
  }

  float pre[kLoadPer];  // p1 of the next plane to stage
  float qo[kQPer], qs[kQPer];  // p0 and src of the next first-step plane
  float os[kOPer];             // src of the next second-step plane
  auto fetch_p1 = [&](int z) {
    const bool zin = z >= 0 && z < nx;
    const float* b = p1 + (zin ? (size_t)z * plane : 0);
#pragma unroll
    for (int j = 0; j < kLoadPer; ++j) pre[j] = (zin && lg[j] >= 0) ? b[lg[j]] : 0.f;
  };
  auto commit_p1 = [&](int z) {
    float* dst = P[z & 3];
#pragma unroll
    for (int j = 0; j < kLoadPer; ++j)
      if (tid + j * kFT < kPW * kPH) dst[tid + j * kFT] = pre[j];
  };
  auto fetch_q = [&](int z) {
    const bool zin = z >= 1 && z <= nx - 2;
    const size_t zb = zin ? (size_t)z * plane : 0;
    const float* b0 = p0 + zb;
    const float* bs = src + zb;
#pragma unroll
    for (int j = 0; j < kQPer; ++j) {
      const bool ok = zin && qg[j] >= 0;
      qo[j] = ok ? b0[qg[j]] : 0.f;
      qs[j] = ok ? bs[qg[j]] : 0.f;
    }
  };
  auto fetch_o = [&](int z) {
    const bool zin = z >= 1 && z <= nx - 2;
    const float* bs = src + (zin ? (size_t)z * plane : 0);
#pragma unroll
    for (int j = 0; j < kOPer; ++j) os[j] = (zin && oi[j] && og[j] >= 0) ? bs[og[j]] : 0.f;
  };

  fetch_p1(z0 - 2);
  commit_p1(z0 - 2);
  fetch_p1(z0 - 1);
  commit_p1(z0 - 1);
  fetch_p1(z0);
  fetch_q(z0 - 1);
  fetch_o(z0);
  // Each prefetch is issued right after its registers were consumed, so
  // the loads are in flight across a barrier and the other step.
  for (int z = z0 - 1; z <= z1; ++z) {
    commit_p1(z + 1);
    fetch_p1(z + 2);
    __syncthreads();
    // first step at plane z over the tile and its one-cell ring
    {
      const float* pc = P[z & 3];
      const float* pm = P[(z - 1) & 3];
      const float* pp = P[(z + 1) & 3];
      float* q = Qs[mod3(z)];
      const bool zint = z >= 1 && z <= nx - 2;
      const float tz = zint ? tab[2 * nx + z] : 0.f;
#pragma unroll
      for (int j = 0; j < kQPer; ++j) {
        const int f = tid + j * kFT;
        if (f >= kQW * kQH) continue;
        const int qy = f / kQW, qx = f - qy * kQW;
        float out = 0.f;
        if (zint && qg[j] >= 0) {
          const int c = (qy + 1) * kPW + qx + 1;
          const float s6 = ((pc[c - 1] + pc[c + 1]) + (pc[c - kPW] + pc[c + kPW])) + (pm[c] + pp[c]);
          out = wave_update(pc[c], s6, qo[j], qs[j], qd[j] + tz, k);
        }
        q[f] = out;
      }
    }
    fetch_q(z + 1);
    __syncthreads();
    // second step at plane z - 1 over the tile
    const int zo = z - 1;
    if (zo >= z0 && zo < z1) {
      const float* qc = Qs[mod3(zo)];
      const float* qm = Qs[mod3(zo - 1)];
      const float* qp = Qs[mod3(z)];
      const float* po = P[zo & 3];
      const bool zint = zo >= 1 && zo <= nx - 2;
      const float tz = zint ? tab[2 * nx + zo] : 0.f;
#pragma unroll
      for (int j = 0; j < kOPer; ++j) {
        if (og[j] < 0) continue;
        const int f = tid + j * kFT;
        const int oy = f / kFX, ox = f - oy * kFX;
        const int c = (oy + 1) * kQW + ox + 1;
        const float cc = qc[c];
        float out = 0.f;
        if (zint && oi[j]) {
          const float s6 = ((qc[c - 1] + qc[c + 1]) + (qc[c - kQW] + qc[c + kQW])) + (qm[c] + qp[c]);
          out = wave_update(cc, s6, po[(oy + 2) * kPW + ox + 2], os[j], od[j] + tz, k);
        }
        // Non-temporal: the fresh fields are not read again in this pass,
        // and kept out of L2 they leave it to the rings that neighbouring
        // tiles re-read (40.3 instead of 41.5 us per iteration at 256^3).
        const size_t at = (size_t)zo * plane + og[j];
        __builtin_nontemporal_store(cc, &p2[at]);
        __builtin_nontemporal_store(out, &p3[at]);
      }
    }
    if (z >= z0) fetch_o(z);  // second step at plane z, next trip
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

void launch_step2(khrr_ctx* ctx, const khrr_wave* w, const WaveConst& k, const float* tab,
                  const float* p0, const float* p1, const float* src, float* p2, float* p3) {
  const long long tiles = (long long)khrr::ceil_div(w->nx, kFX) * khrr::ceil_div(w->nx, kFY) *
                          khrr::ceil_div(w->nx, kFZ);
  const unsigned blocks = (unsigned)((tiles + 7) / 8 * 8);  // see the tile mapping in the kernel
  hipLaunchKernelGGL(wave_step2_kernel, dim3(blocks), dim3(kFT), 0, ctx->stream, w->nx, k, tab, p0,
                     p1, src, p2, p3);
}

// Dissipation tables (3 nx floats) at the start of the scratch arena,
// followed by `extra_floats` more when asked (0 if they do not fit).
int build_tables(khrr_ctx* ctx, const khrr_wave* w, float** tab, size_t extra_floats = 0,
                 float** extra = nullptr) {
  const size_t tb = khrr::Carver::need(3 * (size_t)w->nx, sizeof(float));
  void* base = nullptr;
  if (extra) {
    *extra = nullptr;
    if (khrr::scratch(ctx, tb + khrr::Carver::need(extra_floats, sizeof(float)), &base) == 0)
      *extra = reinterpret_cast<float*>(static_cast<char*>(base) + tb);
    else
      ctx->err.clear();
  }
  if (!base) KHRR_TRY(khrr::scratch(ctx, tb, &base));
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
  const size_t G = (size_t)w->nx * w->nx * w->nx;
  // Pairs of iterations run as one two-step pass, which writes two fresh
  // fields, so a fourth field from the scratch arena joins the caller's
  // three (the one-step kernel is used when it cannot be had).
  float* tab = nullptr;
  float* spare = nullptr;
  KHRR_TRY(build_tables(ctx, w, &tab, iters >= 2 ? G : 0, iters >= 2 ? &spare : nullptr));
  const WaveConst k = make_const(w);
  float* buf[4] = {*prev, *curr, *next, spare};
  int lp = 0, lc = 1;            // buffers holding phi_old and phi
  int fa = 2, fb = spare ? 3 : -1;  // free buffers
  int done = 0;
  {
    khrr::Timed tm(ctx, KHRR_K_WAVE);
    while (fb >= 0 && iters - done >= 2) {
      launch_step2(ctx, w, k, tab, buf[lp], buf[lc], src, buf[fa], buf[fb]);
      const int op = lp, oc = lc;
      lp = fa;
      lc = fb;
      fa = op;
      fb = oc;
      done += 2;
    }
    while (done < iters) {
      launch_step(ctx, w, k, tab, buf[fa], buf[lc], buf[lp], src);
      const int op = lp;
      lp = lc;
      lc = fa;
      fa = op;
      ++done;
    }
    // The reference's rotation (swap(prev, curr); swap(curr, next) per
    // iteration) leaves, after `iters` steps, phi_old in caller buffer
    // tp and phi in tc; move the fields there (at most three copies).
    const int r = iters % 3;
    const int tp = r == 0 ? 0 : (r == 1 ? 1 : 2);
    const int tc = r == 0 ? 1 : (r == 1 ? 2 : 0);
    const size_t bytes = G * sizeof(float);
    while (lp != tp || lc != tc) {
      if (lp != tp && tp != lc) {
        KHRR_HIP(ctx, hipMemcpyAsync(buf[tp], buf[lp], bytes, hipMemcpyDeviceToDevice, ctx->stream));
        lp = tp;
      } else if (lc != tc && tc != lp) {
        KHRR_HIP(ctx, hipMemcpyAsync(buf[tc], buf[lc], bytes, hipMemcpyDeviceToDevice, ctx->stream));
        lc = tc;
      } else {  // tp == lc and tc == lp: park phi_old in a free buffer first
        int f = 0;
        while (f == lp || f == lc || !buf[f]) ++f;
        KHRR_HIP(ctx, hipMemcpyAsync(buf[f], buf[lp], bytes, hipMemcpyDeviceToDevice, ctx->stream));
        lp = f;
      }
    }
  }
  KHRR_TRY(khrr::check_launch(ctx, "wave_step kernels"));
  float* caller[3] = {buf[0], buf[1], buf[2]};
  *prev = caller[lp];
  *curr = caller[lc];
  *next = caller[3 - lp - lc];
  return 0;
}

}  // extern "C"
