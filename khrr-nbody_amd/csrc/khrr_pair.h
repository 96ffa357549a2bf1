// khrr_pair.h -- the softened pair interaction and the j-loop shared by the
// plain and the batched gravity kernels (device code, header-only).
//
// Pair (nbody/include/nbody-kernel.cuh:117-131 for the force, :72-79 for
// the potential):  d = r_j - r_i,  r2 = |d|^2 + h_i + h_j,  inv = rsq(r2),
// phi += m_j inv,  acc += m_j inv^3 d.
#pragma once

#include <hip/hip_runtime.h>

namespace khrr {

constexpr int kJU = 8;  // j unroll (SGPR budget: 5 arrays x 8 values)

__device__ __forceinline__ void fold(float& hi, float& lo, float t, bool comp) {
  if (comp) {
    // TwoSum (Knuth): hi + t = s + e exactly.
    const float s = hi + t;
    const float bb = s - hi;
    const float e = (hi - (s - bb)) + (t - bb);
    hi = s;
    lo += e;
  } else {
    hi += t;
  }
}

template <bool DO_ACC, bool DO_PHI, bool UNIH>
__device__ __forceinline__ void pair(float xj, float yj, float zj, float mj, float hj, float xi,
                                     float yi, float zi, float hh, float& ax, float& ay, float& az,
                                     float& ph) {
  const float dx = xj - xi;
  const float dy = yj - yi;
  const float dz = zj - zi;
  const float h = UNIH ? hh : hh + hj;
  const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, h)));
  const float inv = __builtin_amdgcn_rsqf(r2);
  const float s = mj * inv;
  if (DO_PHI) ph += s;
  if (DO_ACC) {
    const float s3 = s * (inv * inv);
    ax = __builtin_fmaf(dx, s3, ax);
    ay = __builtin_fmaf(dy, s3, ay);
    az = __builtin_fmaf(dz, s3, az);
  }
}

// Pairs of one j-range [j, je) against the IPL i-particles of this lane.
template <int IPL, bool DO_ACC, bool DO_PHI, bool UNIH>
__device__ __forceinline__ void run_pairs(int j, int je, const float* __restrict__ xj,
                                          const float* __restrict__ yj,
                                          const float* __restrict__ zj,
                                          const float* __restrict__ mj,
                                          const float* __restrict__ hj, const float (&xi)[IPL],
                                          const float (&yi)[IPL], const float (&zi)[IPL],
                                          const float (&hh)[IPL], float (&tx)[IPL],
                                          float (&ty)[IPL], float (&tz)[IPL], float (&tp)[IPL]) {
  for (; j + kJU <= je; j += kJU) {
    float sx[kJU], sy[kJU], sz[kJU], sm[kJU], sh[kJU];
#pragma unroll
    for (int u = 0; u < kJU; ++u) {
      sx[u] = xj[j + u];
      sy[u] = yj[j + u];
      sz[u] = zj[j + u];
      sm[u] = mj[j + u];
      sh[u] = UNIH ? 0.f : hj[j + u];
    }
#pragma unroll
    for (int u = 0; u < kJU; ++u) {
#pragma unroll
      for (int k = 0; k < IPL; ++k)
        pair<DO_ACC, DO_PHI, UNIH>(sx[u], sy[u], sz[u], sm[u], sh[u], xi[k], yi[k], zi[k], hh[k],
                                   tx[k], ty[k], tz[k], tp[k]);
    }
  }
  for (; j < je; ++j) {
    const float sx = xj[j], sy = yj[j], sz = zj[j], sm = mj[j];
    const float sh = UNIH ? 0.f : hj[j];
#pragma unroll
    for (int k = 0; k < IPL; ++k)
      pair<DO_ACC, DO_PHI, UNIH>(sx, sy, sz, sm, sh, xi[k], yi[k], zi[k], hh[k], tx[k], ty[k],
                                 tz[k], tp[k]);
  }
}

}  // namespace khrr
