// khrr_diag.hip -- FP64 conservation-law reductions on the device.
//
// Replaces the host loop of result() (nbody/src/nbody.cu:797-833), which
// copies five arrays to the host at every save (:1639-1643) and sums them
// serially.  Here each block reduces a fixed, contiguous slice in a fixed
// tree order and one final block combines the block partials by a second
// fixed tree, so the sums are deterministic run to run (they differ from
// the serial sum only by FP64 reassociation).
#include <cmath>

#include "khrr_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kSums = 12;  // EK EP PX PY PZ LX LY LZ M MX MY MZ
constexpr int kMaxs = 5;   // VR VFI VZ R Z
constexpr int kVals = kSums + kMaxs;
constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;

// One body's state, loaded ahead of its accumulation so that several
// bodies' loads are in flight together.
struct Body {
  double x, y, z, vx, vy, vz, m, p;
};

__device__ __forceinline__ Body load_body(int i, size_t N, const double* __restrict__ pos,
                                          const double* __restrict__ vel,
                                          const double* __restrict__ mass,
                                          const float* __restrict__ phi) {
  Body b;
  b.x = pos[i];
  b.y = pos[N + i];
  b.z = pos[2 * N + i];
  b.vx = vel[i];
  b.vy = vel[N + i];
  b.vz = vel[2 * N + i];
  b.m = mass[i];
  b.p = phi ? (double)phi[i] : 0.0;
  return b;
}

__device__ __forceinline__ void accumulate_one(const Body& b, bool has_phi, double (&v)[kVals]) {
  const double x = b.x, y = b.y, z = b.z, vx = b.vx, vy = b.vy, vz = b.vz, m = b.m;
  const double r = sqrt(x * x + y * y);
  double vr = 0.0, vfi = 0.0;
  if (r > 0.0) {
    vr = (vx * x + vy * y) / r;
    vfi = (vy * x - vx * y) / r;
  }
  v[0] += m * (vx * vx + vy * vy + vz * vz);
  if (has_phi) v[1] += m * b.p;
  v[2] += m * vx;
  v[3] += m * vy;
  v[4] += m * vz;
  v[5] += (vz * y - vy * z) * m;
  v[6] += (vx * z - vz * x) * m;
  v[7] += (vy * x - vx * y) * m;
  v[8] += m;
  v[9] += m * x;
  v[10] += m * y;
  v[11] += m * z;
  v[12] = fmax(v[12], fabs(vr));
  v[13] = fmax(v[13], vfi);
  v[14] = fmax(v[14], fabs(vz));
  v[15] = fmax(v[15], r);
  v[16] = fmax(v[16], fabs(z));
}

// Fixed-order reduction of the block's kVals values through LDS: every
// thread stores its values, then thread 8k + s (s < 8) combines value k of
// threads s, s + 8, ..., s + 8 (kThreads / 8 - 1) in that order, and thread
// k combines those eight results in order s = 0..7.  The order is fixed, so
// the result does not depend on timing.  Value k goes to out[k * stride].
constexpr int kSeg = 8;
static_assert(kVals * kSeg <= kThreads && kThreads % kSeg == 0, "one thread per (value, segment)");
static_assert(kMaxBlocks % kThreads == 0, "whole rows of partials per thread");

__device__ void block_reduce(double (&v)[kVals], double* __restrict__ out, int stride) {
  __shared__ double sh[kVals][kThreads];
  __shared__ double seg[kVals][kSeg];
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < kVals; ++k) sh[k][t] = v[k];
  __syncthreads();
  if (t < kVals * kSeg) {
    const int k = t / kSeg, s = t % kSeg;
    const bool is_max = k >= kSums;
    double x = sh[k][s];
#pragma unroll 8
    for (int j = 1; j < kThreads / kSeg; ++j) {
      const double y = sh[k][j * kSeg + s];
      x = is_max ? fmax(x, y) : x + y;
    }
    seg[k][s] = x;
  }
  __syncthreads();
  if (t < kVals) {
    const bool is_max = t >= kSums;
    double x = seg[t][0];
#pragma unroll
    for (int s = 1; s < kSeg; ++s) x = is_max ? fmax(x, seg[t][s]) : x + seg[t][s];
    out[(size_t)t * stride] = x;
  }
}

// Block b reduces the contiguous slice [b*per, (b+1)*per).  Thread t adds
// bodies t, t + kThreads, t + 2 kThreads, ... in that order; the loads of
// kU consecutive ones are issued before any of them is added.  Value k of
// block b goes to part[k * kMaxBlocks + b].
constexpr int kU = 4;

__global__ __launch_bounds__(kThreads) void diag_partial(int n, int per,
                                                         const double* __restrict__ pos,
                                                         const double* __restrict__ vel,
                                                         const double* __restrict__ mass,
                                                         const float* __restrict__ phi,
                                                         double* __restrict__ part) {
  double v[kVals];
  for (int k = 0; k < kVals; ++k) v[k] = 0.0;
  const int i0 = blockIdx.x * per;
  const int i1 = min(n, i0 + per);
  const bool has_phi = phi != nullptr;
  const int len = i1 - i0;  // offsets within the slice stay far from INT_MAX
  for (int o = threadIdx.x; o < len; o += kU * kThreads) {
    Body b[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (o + u * kThreads < len)
        b[u] = load_body(i0 + o + u * kThreads, (size_t)n, pos, vel, mass, phi);
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (o + u * kThreads < len) accumulate_one(b[u], has_phi, v);
  }
  block_reduce(v, part + blockIdx.x, kMaxBlocks);
}

// Thread t combines blocks t, t + kThreads, ... (nblocks <= kMaxBlocks) in
// that order, reading each value's row of partials with unit stride, then
// the same fixed-order block reduction.
__global__ __launch_bounds__(kThreads) void diag_final(int nblocks, const double* __restrict__ part,
                                                       double* __restrict__ out) {
  constexpr int kRows = kMaxBlocks / kThreads;
  double x[kRows][kVals];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int b = threadIdx.x + r * kThreads;
#pragma unroll
    for (int k = 0; k < kVals; ++k) x[r][k] = b < nblocks ? part[(size_t)k * kMaxBlocks + b] : 0.0;
  }
  double v[kVals];
#pragma unroll
  for (int k = 0; k < kVals; ++k) {
    v[k] = x[0][k];
#pragma unroll
    for (int r = 1; r < kRows; ++r) v[k] = k < kSums ? v[k] + x[r][k] : fmax(v[k], x[r][k]);
  }
  block_reduce(v, out, 1);
}

}  // namespace

namespace khrr {

// Enqueues the reduction; results land in the device array dev_out
// (KHRR_DIAG_COUNT doubles).
int diag_launch(khrr_ctx* ctx, int n, const double* pos, const double* vel, const double* mass,
                const float* phi, double* dev_out, void* scratch_base) {
  int nblocks = (int)ceil_div(n, 4 * kThreads);
  if (nblocks > kMaxBlocks) nblocks = kMaxBlocks;
  if (nblocks < 1) nblocks = 1;
  const int per = (int)ceil_div(n, nblocks);
  double* part = static_cast<double*>(scratch_base);
  Timed tm(ctx, KHRR_K_DIAG);
  hipLaunchKernelGGL(diag_partial, dim3(nblocks), dim3(kThreads), 0, ctx->stream, n, per, pos, vel,
                     mass, phi, part);
  hipLaunchKernelGGL(diag_final, dim3(1), dim3(kThreads), 0, ctx->stream, nblocks, part, dev_out);
  return check_launch(ctx, "diag kernels");
}

size_t diag_scratch_bytes() { return (size_t)kMaxBlocks * kVals * sizeof(double) + 256; }

}  // namespace khrr

extern "C" int khrr_diag(khrr_ctx* ctx, int n, const double* pos, const double* vel,
                         const double* mass, const float* phi, double* out) {
  KHRR_TRY(khrr::enter(ctx));
  if (n < 0 || !out) return khrr::fail(ctx, KHRR_EINVAL, "bad arguments");
  if (n > 0 && (!pos || !vel || !mass)) return khrr::fail(ctx, KHRR_EINVAL, "null state pointer");
  if (n == 0) {
    for (int k = 0; k < KHRR_DIAG_COUNT; ++k) out[k] = 0.0;
    return 0;
  }
  void* base = nullptr;
  const size_t part_bytes = khrr::diag_scratch_bytes();
  KHRR_TRY(khrr::scratch(ctx, part_bytes + 256, &base));
  double* dev_out = reinterpret_cast<double*>(static_cast<char*>(base) + part_bytes);
  KHRR_TRY(khrr::diag_launch(ctx, n, pos, vel, mass, phi, dev_out, base));
  KHRR_HIP(ctx, hipMemcpyAsync(ctx->host_stage, dev_out, KHRR_DIAG_COUNT * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
  KHRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < KHRR_DIAG_COUNT; ++k) out[k] = ctx->host_stage[k];
  return 0;
}
