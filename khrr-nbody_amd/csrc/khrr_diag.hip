// khrr_diag.hip -- FP64 conservation-law reductions on the device.
//
// Replaces the host loop of result() (nbody/src/nbody.cu:797-833), which
// copies five arrays to the host at every save (:1639-1643) and sums them
// serially.  Here each block reduces a fixed, contiguous slice in a fixed
// tree order and one final block combines the block partials in index
// order, so the sums are deterministic run to run (they differ from the
// serial sum only by FP64 reassociation).
#include <cmath>

#include "khrr_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kSums = 12;  // EK EP PX PY PZ LX LY LZ M MX MY MZ
constexpr int kMaxs = 5;   // VR VFI VZ R Z
constexpr int kVals = kSums + kMaxs;
constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;

__device__ __forceinline__ void accumulate_one(int i, size_t N, const double* __restrict__ pos,
                                               const double* __restrict__ vel,
                                               const double* __restrict__ mass,
                                               const float* __restrict__ phi, double (&v)[kVals]) {
  const double x = pos[i], y = pos[N + i], z = pos[2 * N + i];
  const double vx = vel[i], vy = vel[N + i], vz = vel[2 * N + i];
  const double r = sqrt(x * x + y * y);
  double vr = 0.0, vfi = 0.0;
  if (r > 0.0) {
    vr = (vx * x + vy * y) / r;
    vfi = (vy * x - vx * y) / r;
  }
  const double m = mass[i];
  v[0] += m * (vx * vx + vy * vy + vz * vz);
  if (phi) v[1] += m * (double)phi[i];
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

// Fixed-order reduction of the block's kVals values: a shuffle tree inside
// each wave, then the waves' results combined in wave order by one thread
// per value, so the result does not depend on timing.
__device__ void block_reduce(double (&v)[kVals], double* __restrict__ out) {
  constexpr int kWaves = kThreads / khrr::kWave;
  __shared__ double sh[kWaves][kVals];
  const int lane = threadIdx.x & (khrr::kWave - 1);
  const int w = threadIdx.x / khrr::kWave;
#pragma unroll
  for (int k = 0; k < kVals; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = khrr::kWave / 2; off > 0; off >>= 1) {
      const double y = __shfl_down(x, off, khrr::kWave);
      x = k < kSums ? x + y : fmax(x, y);
    }
    if (lane == 0) sh[w][k] = x;
  }
  __syncthreads();
  if ((int)threadIdx.x < kVals) {
    const int k = threadIdx.x;
    double x = sh[0][k];
    for (int q = 1; q < kWaves; ++q) x = k < kSums ? x + sh[q][k] : fmax(x, sh[q][k]);
    out[k] = x;
  }
}

// Block b reduces the contiguous slice [b*per, (b+1)*per).
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
  for (int i = i0 + threadIdx.x; i < i1; i += kThreads)
    accumulate_one(i, (size_t)n, pos, vel, mass, phi, v);
  block_reduce(v, part + (size_t)blockIdx.x * kVals);
}

__global__ __launch_bounds__(kThreads) void diag_final(int nblocks, const double* __restrict__ part,
                                                       double* __restrict__ out) {
  double v[kVals];
  for (int k = 0; k < kVals; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kThreads)
    for (int k = 0; k < kVals; ++k) {
      const double x = part[(size_t)b * kVals + k];
      v[k] = k < kSums ? v[k] + x : fmax(v[k], x);
    }
  block_reduce(v, out);
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
