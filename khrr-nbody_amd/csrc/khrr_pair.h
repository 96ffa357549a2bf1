// khrr_pair.h -- the softened pair interaction and the j-loop shared by the
// plain and the batched gravity kernels (device code, header-only).
//
// Pair (nbody/include/nbody-kernel.cuh:117-131 for the force, :72-79 for
// the potential):  d = r_j - r_i,  r2 = |d|^2 + h_i + h_j,  inv = rsq(r2),
// phi += m_j inv,  acc += m_j inv^3 d.
//
// A j-range whose softening (h) or mass is the same for every body is
// flagged; then h_i + h_j is formed once per i instead of per pair, and the
// mass multiplies the range's partial sum once instead of every pair term.
// A pair costs 11 VALU ops + one v_rsq_f32 with both flags, 13 + 1 with none.
#pragma once

#include <hip/hip_runtime.h>

namespace khrr {

constexpr int kJU = 8;        // j unroll (SGPR budget: 5 arrays x 8 values)
constexpr int kUniformH = 1;  // every h_j of the range is equal
constexpr int kUniformM = 2;  // every m_j of the range is equal

// TwoSum with the error term kept in LDS (lo: this lane's slot), for
// kernels whose registers cannot also hold the error terms.
__device__ __forceinline__ void fold_lds(float& hi, float* lo, float t) {
  const float s = hi + t;
  const float bb = s - hi;
  const float e = (hi - (s - bb)) + (t - bb);
  hi = s;
  *lo += e;
}

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

template <bool DO_ACC, bool DO_PHI, bool UNIH, bool UNIM>
__device__ __forceinline__ void pair(float xj, float yj, float zj, float mj, float hj, float xi,
                                     float yi, float zi, float hh, float& ax, float& ay, float& az,
                                     float& ph) {
  const float dx = xj - xi;
  const float dy = yj - yi;
  const float dz = zj - zi;
  const float h = UNIH ? hh : hh + hj;
  const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, h)));
  const float inv = __builtin_amdgcn_rsqf(r2);
  const float s = UNIM ? inv : mj * inv;
  if (DO_PHI) ph += s;
  if (DO_ACC) {
    const float s3 = s * (inv * inv);
    ax = __builtin_fmaf(dx, s3, ax);
    ay = __builtin_fmaf(dy, s3, ay);
    az = __builtin_fmaf(dz, s3, az);
  }
}

// Pairs of one j-range [j, je) against the IPL i-particles of this lane.
// j is wave-uniform, so every j load is a scalar load.
template <int IPL, bool DO_ACC, bool DO_PHI, bool UNIH, bool UNIM>
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
      sm[u] = UNIM ? 0.f : mj[j + u];
      sh[u] = UNIH ? 0.f : hj[j + u];
    }
#pragma unroll
    for (int u = 0; u < kJU; ++u) {
#pragma unroll
      for (int k = 0; k < IPL; ++k)
        pair<DO_ACC, DO_PHI, UNIH, UNIM>(sx[u], sy[u], sz[u], sm[u], sh[u], xi[k], yi[k], zi[k],
                                         hh[k], tx[k], ty[k], tz[k], tp[k]);
    }
  }
  for (; j < je; ++j) {
    const float sx = xj[j], sy = yj[j], sz = zj[j];
    const float sm = UNIM ? 0.f : mj[j];
    const float sh = UNIH ? 0.f : hj[j];
#pragma unroll
    for (int k = 0; k < IPL; ++k)
      pair<DO_ACC, DO_PHI, UNIH, UNIM>(sx, sy, sz, sm, sh, xi[k], yi[k], zi[k], hh[k], tx[k],
                                       ty[k], tz[k], tp[k]);
  }
}

// Per-lane accumulators of IPL i-particles: running sums (a*) and, for the
// compensated form, their TwoSum error terms (l*).
template <int IPL>
struct Accum {
  float ax[IPL], ay[IPL], az[IPL], ap[IPL];
  float lx[IPL], ly[IPL], lz[IPL], lp[IPL];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
      ax[k] = ay[k] = az[k] = ap[k] = 0.f;
      lx[k] = ly[k] = lz[k] = lp[k] = 0.f;
    }
  }
};

// One j-tile [jt, je): pairs into fresh FP32 partials, then one fold of the
// partials into the running sums.  `flags` (kUniformH | kUniformM, wave
// uniform) selects the specialised loop.  lo (compensated form only): the
// error terms in LDS instead of acc.l*, this lane's slot of a
// [components][IPL][64 lanes] block (ax, ay, az, then the potential; the
// potential alone in row 0).
template <int IPL, bool DO_ACC, bool DO_PHI, bool COMP>
__device__ __forceinline__ void tile_step(int jt, int je, int flags, const float* __restrict__ xj,
                                          const float* __restrict__ yj,
                                          const float* __restrict__ zj,
                                          const float* __restrict__ mj,
                                          const float* __restrict__ hj, const float (&xi)[IPL],
                                          const float (&yi)[IPL], const float (&zi)[IPL],
                                          const float (&hi)[IPL], Accum<IPL>& acc,
                                          float* lo = nullptr) {
  float tx[IPL], ty[IPL], tz[IPL], tp[IPL];
#pragma unroll
  for (int k = 0; k < IPL; ++k) tx[k] = ty[k] = tz[k] = tp[k] = 0.f;
  const bool uh = (flags & kUniformH) != 0;
  const bool um = (flags & kUniformM) != 0;
  if (uh) {
    const float h0 = hj[jt];
    float hh[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) hh[k] = hi[k] + h0;
    if (um)
      run_pairs<IPL, DO_ACC, DO_PHI, true, true>(jt, je, xj, yj, zj, mj, hj, xi, yi, zi, hh, tx, ty,
                                                 tz, tp);
    else
      run_pairs<IPL, DO_ACC, DO_PHI, true, false>(jt, je, xj, yj, zj, mj, hj, xi, yi, zi, hh, tx,
                                                  ty, tz, tp);
  } else {
    if (um)
      run_pairs<IPL, DO_ACC, DO_PHI, false, true>(jt, je, xj, yj, zj, mj, hj, xi, yi, zi, hi, tx,
                                                  ty, tz, tp);
    else
      run_pairs<IPL, DO_ACC, DO_PHI, false, false>(jt, je, xj, yj, zj, mj, hj, xi, yi, zi, hi, tx,
                                                   ty, tz, tp);
  }
  const float scale = um ? mj[jt] : 1.f;
#pragma unroll
  for (int k = 0; k < IPL; ++k) {
    if (COMP && lo) {
      if (DO_ACC) {
        fold_lds(acc.ax[k], lo + (0 * IPL + k) * 64, tx[k] * scale);
        fold_lds(acc.ay[k], lo + (1 * IPL + k) * 64, ty[k] * scale);
        fold_lds(acc.az[k], lo + (2 * IPL + k) * 64, tz[k] * scale);
      }
      if (DO_PHI) fold_lds(acc.ap[k], lo + ((DO_ACC ? 3 : 0) * IPL + k) * 64, tp[k] * scale);
      continue;
    }
    if (DO_ACC) {
      fold(acc.ax[k], acc.lx[k], tx[k] * scale, COMP);
      fold(acc.ay[k], acc.ly[k], ty[k] * scale, COMP);
      fold(acc.az[k], acc.lz[k], tz[k] * scale, COMP);
    }
    if (DO_PHI) fold(acc.ap[k], acc.lp[k], tp[k] * scale, COMP);
  }
}

// Final outputs of lane-particle k (global index i, planes of stride n):
// acc (+)= sum, phi (-)= sum (ACCEL_kernel :137-141, PHI_kernel :85).
template <int IPL, bool DO_ACC, bool DO_PHI, bool COMP>
__device__ __forceinline__ void store_final(const Accum<IPL>& a, int k, int i, size_t n,
                                            float* __restrict__ acc, float* __restrict__ phi,
                                            int accumulate) {
  if (DO_ACC) {
    const float fx = COMP ? a.ax[k] + a.lx[k] : a.ax[k];
    const float fy = COMP ? a.ay[k] + a.ly[k] : a.ay[k];
    const float fz = COMP ? a.az[k] + a.lz[k] : a.az[k];
    float* p0 = acc + i;
    float* p1 = acc + n + i;
    float* p2 = acc + 2 * n + i;
    if (accumulate) {
      *p0 = *p0 + fx;
      *p1 = *p1 + fy;
      *p2 = *p2 + fz;
    } else {
      *p0 = fx;
      *p1 = fy;
      *p2 = fz;
    }
  }
  if (DO_PHI) {
    const float fp = COMP ? a.ap[k] + a.lp[k] : a.ap[k];
    phi[i] = accumulate ? phi[i] - fp : -fp;
  }
}

}  // namespace khrr
