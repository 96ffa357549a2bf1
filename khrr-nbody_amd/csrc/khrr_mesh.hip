// khrr_mesh.hip -- mesh force path of the wave solver: the acceleration
// field of the stencil's potential and its transfer to the particles.
//
// Replaces acc_field (nbody/include/p2mesh-kernel.cuh:169-224),
// apply_acceleration_to_particles (:226-238) and apply_cell_to_particles
// (:240-252), called by calc_acceleration_by_grid (nbody/src/nbody.cu:243-262)
// after the wave iterations, and acc_abs (:289-298).  The FP64 form of the
// field and acc_abs serve the debug dumps of convert_particles
// (nbody.cu:287-701), which difference the FP64 cell potential of the
// bodies and take |a| of body and cell accelerations.
//
// acc_field indexes cells in 3-D but the reference launches it over a 1-D
// grid of cells (nbody.cu:244, over_cells x over_blocks), so there every
// cell is evaluated as if it sat at iy = iz = 0 (one-sided y and z
// differences against index + NX and + NX^2, central x differences across
// row ends) and the last NX^2 cells read past the field.  Here every cell
// gets the stencil its (ix, iy, iz) position asks for: central differences
// inside, one-sided at the faces, times -1/DX.
//
// Arithmetic: the FP32 field values are widened to FP64, the reference's
// expression is evaluated in FP64 (the difference of two floats is exact
// there), and the result is rounded to FP32 once; the FP64 form is the
// same expression on FP64 values.  HBM traffic per cell:
// read phi (4 B, neighbours from cache), write 3 x 4 B (+4 B for |a|).
#include <cmath>

#include "khrr_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kMX = 64, kMY = 4, kMZ = 32;

__device__ __forceinline__ double diff1(bool lo_face, bool hi_face, double minus, double c,
                                        double plus) {
  if (lo_face) return plus - c;
  if (hi_face) return c - minus;
  return 0.5 * (plus - minus);
}

// T = float: the field is widened to FP64 on load and the results rounded
// to FP32 once; T = double: plain FP64.
template <typename T>
__global__ __launch_bounds__(kMX* kMY) void mesh_acc_kernel(int nx, double coef,
                                                            const T* __restrict__ phi,
                                                            T* __restrict__ acc,
                                                            T* __restrict__ acc_abs) {
  const int ix = blockIdx.x * kMX + threadIdx.x;
  const int iy = blockIdx.y * kMY + threadIdx.y;
  if (ix >= nx || iy >= nx) return;
  const int z0 = blockIdx.z * kMZ;
  const int z1 = min(nx, z0 + kMZ);
  const size_t plane = (size_t)nx * nx;
  const size_t G = plane * nx;
  size_t at = (size_t)ix + (size_t)iy * nx + (size_t)z0 * plane;
  double cm = z0 > 0 ? (double)phi[at - plane] : 0.0;
  double cc = phi[at];
  for (int z = z0; z < z1; ++z, at += plane) {
    const double cp = z + 1 < nx ? (double)phi[at + plane] : 0.0;
    const double xm = ix > 0 ? (double)phi[at - 1] : 0.0;
    const double xp = ix + 1 < nx ? (double)phi[at + 1] : 0.0;
    const double ym = iy > 0 ? (double)phi[at - nx] : 0.0;
    const double yp = iy + 1 < nx ? (double)phi[at + nx] : 0.0;
    double dx = diff1(ix == 0, ix == nx - 1, xm, cc, xp);
    double dy = diff1(iy == 0, iy == nx - 1, ym, cc, yp);
    double dz = diff1(z == 0, z == nx - 1, cm, cc, cp);
    dx *= coef;
    dy *= coef;
    dz *= coef;
    // Non-temporal: the outputs are not read again here, and kept out of
    // L2 they leave it to the field planes the neighbouring columns re-read.
    __builtin_nontemporal_store((T)dx, &acc[at]);
    __builtin_nontemporal_store((T)dy, &acc[G + at]);
    __builtin_nontemporal_store((T)dz, &acc[2 * G + at]);
    if (acc_abs) __builtin_nontemporal_store((T)sqrt(dx * dx + dy * dy + dz * dz), &acc_abs[at]);
    cm = cc;
    cc = cp;
  }
}

// acc_abs (p2mesh-kernel.cuh:289-298): |v| of FP32 vec3 planes, in FP64
// (norm3, common.cuh:84-96: (x*x + y*y) + z*z, then sqrt).
__global__ __launch_bounds__(256) void acc_abs_kernel(int n, const float* __restrict__ v,
                                                      double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t N = (size_t)n;
  const double x = v[i], y = v[N + i], z = v[2 * N + i];
  out[i] = sqrt(x * x + y * y + z * z);
}

// acc[c][i] = cell_acc[c][cell_id[i]], phi[i] = cell_phi[cell_id[i]]; an
// index outside [0, G) gives NaN.
__global__ __launch_bounds__(256) void mesh_gather_kernel(int n, long long G,
                                                          const int* __restrict__ cell_id,
                                                          const float* __restrict__ cell_acc,
                                                          const float* __restrict__ cell_phi,
                                                          float* __restrict__ acc,
                                                          float* __restrict__ phi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cell_id[i];
  const bool ok = c >= 0 && c < G;
  const float nan = __builtin_nanf("");
  if (acc) {
    const size_t N = (size_t)n;
    acc[i] = ok ? cell_acc[c] : nan;
    acc[N + i] = ok ? cell_acc[G + c] : nan;
    acc[2 * N + i] = ok ? cell_acc[2 * G + c] : nan;
  }
  if (phi) phi[i] = ok ? cell_phi[c] : nan;
}

}  // namespace

template <typename T>
static int mesh_acc(khrr_ctx* ctx, int nx, double dx, const T* cell_phi, T* cell_acc,
                    T* cell_acc_abs) {
  KHRR_TRY(khrr::enter(ctx));
  if (nx < 2 || (long long)nx * nx * nx > (1ll << 31))
    return khrr::fail(ctx, KHRR_EINVAL, "nx must be in [2, 1290] (got %d)", nx);
  if (!(dx > 0.0)) return khrr::fail(ctx, KHRR_EINVAL, "dx must be > 0");
  if (!cell_phi || !cell_acc) return khrr::fail(ctx, KHRR_EINVAL, "null field pointer");
  const double coef = -1. / dx;  // p2mesh-kernel.cuh:215
  const dim3 block(kMX, kMY);
  const dim3 grid(khrr::ceil_div(nx, kMX), khrr::ceil_div(nx, kMY), khrr::ceil_div(nx, kMZ));
  khrr::Timed tm(ctx, KHRR_K_WAVE);
  hipLaunchKernelGGL(mesh_acc_kernel<T>, grid, block, 0, ctx->stream, nx, coef, cell_phi,
                     cell_acc, cell_acc_abs);
  return khrr::check_launch(ctx, "mesh_acc_kernel");
}

extern "C" {

int khrr_mesh_acc(khrr_ctx* ctx, int nx, double dx, const float* cell_phi, float* cell_acc,
                  float* cell_acc_abs) {
  return mesh_acc<float>(ctx, nx, dx, cell_phi, cell_acc, cell_acc_abs);
}

int khrr_mesh_acc_f64(khrr_ctx* ctx, int nx, double dx, const double* cell_phi,
                      double* cell_acc, double* cell_acc_abs) {
  return mesh_acc<double>(ctx, nx, dx, cell_phi, cell_acc, cell_acc_abs);
}

int khrr_acc_abs(khrr_ctx* ctx, int n, const float* v, double* out) {
  KHRR_TRY(khrr::enter(ctx));
  if (n < 0) return khrr::fail(ctx, KHRR_EINVAL, "bad size n=%d", n);
  if (n > 0 && (!v || !out)) return khrr::fail(ctx, KHRR_EINVAL, "null pointer");
  if (n == 0) return 0;
  khrr::Timed tm(ctx, KHRR_K_BIN);
  hipLaunchKernelGGL(acc_abs_kernel, dim3(khrr::ceil_div(n, 256)), dim3(256), 0, ctx->stream, n,
                     v, out);
  return khrr::check_launch(ctx, "acc_abs_kernel");
}

int khrr_mesh_to_particles(khrr_ctx* ctx, int n, long long n_cells, const int* cell_id,
                           const float* cell_acc, const float* cell_phi, float* acc, float* phi) {
  KHRR_TRY(khrr::enter(ctx));
  if (n < 0 || n_cells < 1 || n_cells > INT32_MAX)
    return khrr::fail(ctx, KHRR_EINVAL, "bad sizes n=%d cells=%lld", n, n_cells);
  if (!acc && !phi) return khrr::fail(ctx, KHRR_EINVAL, "acc and phi are both NULL");
  if (n > 0 && (!cell_id || (acc && !cell_acc) || (phi && !cell_phi)))
    return khrr::fail(ctx, KHRR_EINVAL, "null input pointer");
  if (n == 0) return 0;
  khrr::Timed tm(ctx, KHRR_K_BIN);
  hipLaunchKernelGGL(mesh_gather_kernel, dim3(khrr::ceil_div(n, 256)), dim3(256), 0, ctx->stream,
                     n, n_cells, cell_id, cell_acc, cell_phi, acc, phi);
  return khrr::check_launch(ctx, "mesh_gather_kernel");
}

}  // extern "C"
