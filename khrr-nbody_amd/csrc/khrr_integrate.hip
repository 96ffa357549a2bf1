// khrr_integrate.hip -- predictor/corrector step and the external halo +
// bulge force, FP64 state.
//
// Replaces kernelNbody_integTime (nbody/include/nbody-kernel.cuh:144-194)
// and dev_fex (:8-40); constants as nbody.cu:1111-1120.  The arithmetic is
// the reference's, operation for operation, with FMA contraction off so
// the result is bit-identical to an IEEE FP64 evaluation of the same
// expressions (dev_fex's atan/log come from the device math library).
// The state is SoA planes, so every access is a coalesced 8-byte stream;
// the kernel is HBM-bound (60 B/particle predictor, 84 B corrector).
#include <cmath>

#include "khrr_internal.h"

#pragma clang fp contract(off)

namespace {

struct ExtPot {
  double a, b, Rb, Rh2, con, const1, Mh_inf, Mb;
};

__device__ __forceinline__ void dev_fex(const ExtPot& e, double px, double py, double pz,
                                        double& fx, double& fy, double& fz) {
  const double rr = sqrt(px * px + py * py + pz * pz);
  if (rr > 0.0) {
    const double rr3 = rr * rr * rr;
    const double rrcore = rr / e.a;
    const double rrbcore1 = rr / e.b;
    const double root1 = sqrt(1.0 + rrbcore1 * rrbcore1);
    double forcehalo, forceblg1;
    if (rr < e.Rh2)
      forcehalo = -e.con * (rrcore - atan(rrcore)) / rr3;
    else
      forcehalo = -e.Mh_inf / rr3;
    if (rr < e.Rb)
      forceblg1 = -e.const1 * (e.b * log(rrbcore1 + root1) - rr / root1) / rr3;
    else
      forceblg1 = -e.Mb / rr3;
    const double forcesph = forceblg1 + forcehalo;
    fx = forcesph * px;
    fy = forcesph * py;
    fz = forcesph * pz;
  } else {
    fx = fy = fz = 0.0;
  }
}

// One body of the predictor (istep 0) or corrector (istep 1).  N is the
// plane stride of the FP64 arrays, L that of the FP32 arrays.
template <bool EXT>
__device__ __forceinline__ void integrate_body(
    int i, size_t N, size_t L, int istep, double dt, const ExtPot& ext,
    const float* __restrict__ acc, double* __restrict__ pos_t, double* __restrict__ vel_t,
    const double* __restrict__ pos, const double* __restrict__ vel, float* __restrict__ acc0,
    float* __restrict__ pos32) {
  const double rx = pos[i], ry = pos[N + i], rz = pos[2 * N + i];
  const double vx = vel[i], vy = vel[N + i], vz = vel[2 * N + i];
  double ex = 0.0, ey = 0.0, ez = 0.0;
  if (EXT) dev_fex(ext, rx, ry, rz, ex, ey, ez);
  const float a_x = acc[i], a_y = acc[L + i], a_z = acc[2 * L + i];
  const double fx = (double)a_x + ex;
  const double fy = (double)a_y + ey;
  const double fz = (double)a_z + ez;
  if (istep == 0) {
    const double tx = vx + dt * fx;
    const double ty = vy + dt * fy;
    const double tz = vz + dt * fz;
    const double px = rx + 0.5 * dt * (vx + tx);
    const double py = ry + 0.5 * dt * (vy + ty);
    const double pz = rz + 0.5 * dt * (vz + tz);
    pos_t[i] = px;
    pos_t[N + i] = py;
    pos_t[2 * N + i] = pz;
    vel_t[i] = tx;
    vel_t[N + i] = ty;
    vel_t[2 * N + i] = tz;
    if (pos32) {
      pos32[i] = (float)px;
      pos32[L + i] = (float)py;
      pos32[2 * L + i] = (float)pz;
    }
  } else {
    const double tx = vel_t[i], ty = vel_t[N + i], tz = vel_t[2 * N + i];
    const double ux = 0.5 * (tx + vx + dt * fx);
    const double uy = 0.5 * (ty + vy + dt * fy);
    const double uz = 0.5 * (tz + vz + dt * fz);
    pos_t[i] = rx;
    pos_t[N + i] = ry;
    pos_t[2 * N + i] = rz;
    vel_t[i] = ux;
    vel_t[N + i] = uy;
    vel_t[2 * N + i] = uz;
    if (acc0) {
      acc0[i] = a_x;
      acc0[L + i] = a_y;
      acc0[2 * L + i] = a_z;
    }
    if (pos32) {
      pos32[i] = (float)rx;
      pos32[L + i] = (float)ry;
      pos32[2 * L + i] = (float)rz;
    }
  }
}

template <bool EXT>
__global__ __launch_bounds__(256) void integrate_kernel(
    int n, int istep, double dt, ExtPot ext, const float* __restrict__ acc,
    double* __restrict__ pos_t, double* __restrict__ vel_t, const double* __restrict__ pos,
    const double* __restrict__ vel, float* __restrict__ acc0, float* __restrict__ pos32,
    int ld32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  integrate_body<EXT>(i, (size_t)n, (size_t)ld32, istep, dt, ext, acc, pos_t, vel_t, pos, vel,
                      acc0, pos32);
}

// Many systems in one launch: block b belongs to the last segment whose
// first_block <= b.  Each segment's FP64 state is one allocation of planes
// pos, vel, pos_t, vel_t (3n each); the FP32 arrays are shared planes of
// stride ld32 with the segment at `begin`.  istep 0 reads pos/vel and
// writes pos_t/vel_t; istep 1 reads pos_t/vel_t and writes pos/vel, exactly
// as the single-system predictor and corrector calls.
__global__ __launch_bounds__(256) void integrate_segments_kernel(
    int n_seg, const khrr::IntSeg* __restrict__ segs, int istep, const float* __restrict__ acc,
    float* __restrict__ acc0, float* __restrict__ pos32, int ld32) {
  int lo = 0, hi = n_seg - 1;
  const int b = blockIdx.x;
  while (lo < hi) {  // uniform per block
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].first_block <= b) lo = mid; else hi = mid - 1;
  }
  const khrr::IntSeg sg = segs[lo];
  const int i = (b - sg.first_block) * blockDim.x + threadIdx.x;
  if (i >= sg.n) return;
  const size_t N = (size_t)sg.n;
  double* P = sg.state;
  double* V = sg.state + 3 * N;
  double* PT = sg.state + 6 * N;
  double* VT = sg.state + 9 * N;
  const size_t off = (size_t)sg.begin;
  const ExtPot none{};
  if (istep == 0)
    integrate_body<false>(i, N, (size_t)ld32, 0, sg.dt, none, acc + off, PT, VT, P, V, nullptr,
                          pos32 ? pos32 + off : nullptr);
  else
    integrate_body<false>(i, N, (size_t)ld32, 1, sg.dt, none, acc + off, P, V, PT, VT,
                          acc0 ? acc0 + off : nullptr, pos32 ? pos32 + off : nullptr);
}

}  // namespace

namespace khrr {

int integrate_launch(khrr_ctx* ctx, int n, int istep, double dt, const khrr_extpot* ext,
                     const float* acc, double* pos_t, double* vel_t, const double* pos,
                     const double* vel, float* acc0, float* pos32, int ld32) {
  if (n == 0) return 0;
  ExtPot e{};
  const bool has_ext = ext && (ext->Mh != 0.0 || ext->Mb != 0.0);
  if (ext) {
    e.a = ext->a;
    e.b = ext->b;
    e.Rb = ext->Rb;
    e.Rh2 = ext->Rh2;
    e.con = ext->con;
    e.const1 = ext->const1;
    e.Mh_inf = ext->Mh_inf;
    e.Mb = ext->Mb;
  }
  Timed tm(ctx, KHRR_K_INTEGRATE);
  const dim3 grid(ceil_div(n, 256)), block(256);
  if (has_ext)
    hipLaunchKernelGGL(integrate_kernel<true>, grid, block, 0, ctx->stream, n, istep, dt, e, acc,
                       pos_t, vel_t, pos, vel, acc0, pos32, ld32);
  else
    hipLaunchKernelGGL(integrate_kernel<false>, grid, block, 0, ctx->stream, n, istep, dt, e, acc,
                       pos_t, vel_t, pos, vel, acc0, pos32, ld32);
  return check_launch(ctx, "integrate_kernel");
}

int integrate_segments(khrr_ctx* ctx, int n_seg, const IntSeg* segs, int total_blocks, int istep,
                       const float* acc, float* acc0, float* pos32, int ld32) {
  if (n_seg == 0 || total_blocks == 0) return 0;
  Timed tm(ctx, KHRR_K_INTEGRATE);
  hipLaunchKernelGGL(integrate_segments_kernel, dim3(total_blocks), dim3(256), 0, ctx->stream,
                     n_seg, segs, istep, acc, acc0, pos32, ld32);
  return check_launch(ctx, "integrate_segments_kernel");
}

}  // namespace khrr

extern "C" {

int khrr_extpot_init(double Mh, double a, double Rh, double Mb, double b, double Rb,
                     khrr_extpot* out) {
  if (!out) return khrr::fail(nullptr, KHRR_EINVAL, "out is null");
  // nbody.cu:1111-1120
  const double Rh2 = 3.0 * Rh;
  const double rcore1 = Rh / a;
  const double rbcore1 = 1.0 / b;
  const double rbcore2 = rbcore1 * rbcore1;
  const double root1 = std::sqrt(1.0 + (Rb * Rb) * rbcore2);
  const double con = Mh / (rcore1 - std::atan(rcore1));
  const double const1 = Mb / (b * std::log(Rb * rbcore1 + root1) - Rb / root1);
  out->Mh = Mh;
  out->a = a;
  out->Rh = Rh;
  out->Mb = Mb;
  out->b = b;
  out->Rb = Rb;
  out->Rh2 = Rh2;
  out->con = con;
  out->const1 = const1;
  out->c_phi_h = con / a * (0.5 * std::log(Rh2 * Rh2 / a / a + 1.0) + std::atan(Rh2 / a) * a / Rh2) +
                 Mh / Rh2;
  out->c_phi_b = Mb / Rb - const1 * std::log(Rb / b + root1) / Rb;
  out->Mh_inf = Mh * (Rh2 / a - std::atan(Rh2 / a)) / (rcore1 - std::atan(rcore1));
  return 0;
}

int khrr_extpot_force(const khrr_extpot* ext, const double p[3], double f[3]) {
  if (!ext || !p || !f) return khrr::fail(nullptr, KHRR_EINVAL, "null argument");
  const double rr = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  if (rr > 0.0) {
    const double rr3 = rr * rr * rr;
    const double rrcore = rr / ext->a;
    const double rrbcore1 = rr / ext->b;
    const double root1 = std::sqrt(1.0 + rrbcore1 * rrbcore1);
    const double fh = rr < ext->Rh2 ? -ext->con * (rrcore - std::atan(rrcore)) / rr3
                                    : -ext->Mh_inf / rr3;
    const double fb = rr < ext->Rb
                          ? -ext->const1 * (ext->b * std::log(rrbcore1 + root1) - rr / root1) / rr3
                          : -ext->Mb / rr3;
    const double fs = fb + fh;
    for (int k = 0; k < 3; ++k) f[k] = fs * p[k];
  } else {
    f[0] = f[1] = f[2] = 0.0;
  }
  return 0;
}

int khrr_integrate(khrr_ctx* ctx, int n, int istep, double dt, double t, const khrr_extpot* ext,
                   const float* acc, double* pos_t, double* vel_t, const double* pos,
                   const double* vel, float* acc0, float* pos32) {
  (void)t;
  KHRR_TRY(khrr::enter(ctx));
  if (n < 0) return khrr::fail(ctx, KHRR_EINVAL, "n < 0");
  if (istep != 0 && istep != 1) return khrr::fail(ctx, KHRR_EINVAL, "istep must be 0 or 1");
  if (n > 0 && (!acc || !pos_t || !vel_t || !pos || !vel))
    return khrr::fail(ctx, KHRR_EINVAL, "null state pointer");
  return khrr::integrate_launch(ctx, n, istep, dt, ext, acc, pos_t, vel_t, pos, vel, acc0, pos32,
                                n);
}

}  // extern "C"
