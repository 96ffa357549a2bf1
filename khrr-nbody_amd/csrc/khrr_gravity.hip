// khrr_gravity.hip -- softened direct-sum gravity on the FP32 VALU (gfx950).
//
// Replaces ACCEL_kernel (nbody/include/nbody-kernel.cuh:89-142) and
// PHI_kernel (:43-86).  Per pair (i, j):
//   d = r_j - r_i;  r2 = |d|^2 + h_i + h_j;  inv = rsq(r2)
//   phi_i -= m_j inv;            acc_i += m_j inv^3 d
// (h = eps^2/2, so h_i + h_j is the reference's 0.5*(eps2_i + eps2_j)).
//
// Design (CDNA4, wave64):
//  * Each wave owns IPL*64 i-particles (IPL per lane, lane-contiguous so the
//    i loads coalesce) and one j-range.  j-particle data is identical for
//    every lane, so it is read with wave-uniform addresses: the compiler
//    emits scalar loads into SGPRs and every VALU op takes the j value as
//    its one scalar operand.  No LDS, no barriers; waves run independently.
//  * The pair is 12-13 VALU ops + one v_rsq_f32: 3 sub, 3 fma for r2 (+1
//    add for h_i + h_j unless the j-tile has uniform softening), 3 mul,
//    3 fma accumulate.  Packed FP32 is disabled at build time, so the
//    issue ceiling is one FP32 op per lane per clock.
//  * Sums are formed per j-tile (j_tile pairs) in FP32 partials and folded
//    into the running total once per tile, either by a plain add or by
//    TwoSum (KHRR_COMPENSATED).  The tile partial keeps the long running
//    sum from absorbing the rounding of every pair term.
//  * j_split > 1 gives several waves per i-group on disjoint j-ranges; the
//    partials are combined in fixed order by a second kernel, so results
//    are deterministic for a given launch shape.
#include "khrr_internal.h"
#include "khrr_pair.h"

namespace {

using khrr::Accum;
using khrr::fold;

// One wave = (i-group, j-range).  Writes either the final outputs (part ==
// nullptr) or its partial sums part[(s*NC + c)*n_i + i], NC = 4 (8 when
// compensated: hi then lo).
template <int IPL, bool DO_ACC, bool DO_PHI, bool COMP>
__global__ __launch_bounds__(256) void grav_kernel(
    int n_i, const float* __restrict__ pos_i, const float* __restrict__ h_i, int n_j,
    const float* __restrict__ pos_j, const float* __restrict__ m_j,
    const float* __restrict__ h_j, const int* __restrict__ tile_flags, int groups_i,
    int n_split, int j_chunk, int j_tile, float* __restrict__ acc, float* __restrict__ phi,
    float* __restrict__ part, int accumulate) {
  const int lane = threadIdx.x & (khrr::kWave - 1);
  // Wave index made provably uniform so every j address below is scalar.
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x / khrr::kWave);
  const int w = blockIdx.x * (blockDim.x / khrr::kWave) + wib;
  if (w >= groups_i * n_split) return;
  const int s = w / groups_i;
  const int ig = w - s * groups_i;
  const int j0 = s * j_chunk;
  const int j1 = min(n_j, j0 + j_chunk);

  const float* xj = pos_j;
  const float* yj = pos_j + n_j;
  const float* zj = pos_j + 2 * (size_t)n_j;

  float xi[IPL], yi[IPL], zi[IPL], hi[IPL];
  int idx[IPL];
#pragma unroll
  for (int k = 0; k < IPL; ++k) {
    idx[k] = ig * (IPL * khrr::kWave) + k * khrr::kWave + lane;
    const bool ok = idx[k] < n_i;
    xi[k] = ok ? pos_i[idx[k]] : 0.f;
    yi[k] = ok ? pos_i[n_i + idx[k]] : 0.f;
    zi[k] = ok ? pos_i[2 * (size_t)n_i + idx[k]] : 0.f;
    hi[k] = ok ? h_i[idx[k]] : 1.f;
  }

  Accum<IPL> a;
  a.zero();
  for (int jt = j0; jt < j1; jt += j_tile) {
    const int je = min(j1, jt + j_tile);
    khrr::tile_step<IPL, DO_ACC, DO_PHI, COMP>(jt, je, tile_flags[jt / j_tile], xj, yj, zj, m_j,
                                               h_j, xi, yi, zi, hi, a);
  }

#pragma unroll
  for (int k = 0; k < IPL; ++k) {
    const int i = idx[k];
    if (i < n_i && part) {
      constexpr int NC = COMP ? 8 : 4;
      float* p = part + (size_t)s * NC * n_i;
      if (DO_ACC) {
        p[i] = a.ax[k];
        p[(size_t)n_i + i] = a.ay[k];
        p[2 * (size_t)n_i + i] = a.az[k];
      }
      if (DO_PHI) p[3 * (size_t)n_i + i] = a.ap[k];
      if (COMP) {
        if (DO_ACC) {
          p[4 * (size_t)n_i + i] = a.lx[k];
          p[5 * (size_t)n_i + i] = a.ly[k];
          p[6 * (size_t)n_i + i] = a.lz[k];
        }
        if (DO_PHI) p[7 * (size_t)n_i + i] = a.lp[k];
      }
    } else if (i < n_i) {
      khrr::store_final<IPL, DO_ACC, DO_PHI, COMP>(a, k, i, (size_t)n_i, acc, phi, accumulate);
    }
  }
}

// Combines the j_split partial sums in split order.
template <bool COMP>
__global__ __launch_bounds__(256) void grav_reduce(int n_i, int n_split, int do_acc, int do_phi,
                                                   const float* __restrict__ part,
                                                   float* __restrict__ acc,
                                                   float* __restrict__ phi, int accumulate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_i) return;
  constexpr int NC = COMP ? 8 : 4;
  const size_t N = (size_t)n_i;
  for (int c = 0; c < 4; ++c) {
    if (c < 3 && !do_acc) continue;
    if (c == 3 && !do_phi) continue;
    float hi = part[c * N + i];
    float lo = COMP ? part[(c + 4) * N + i] : 0.f;
    for (int s = 1; s < n_split; ++s) {
      const float* p = part + (size_t)s * NC * N;
      fold(hi, lo, p[c * N + i], COMP);
      if (COMP) lo += p[(c + 4) * N + i];
    }
    const float v = COMP ? hi + lo : hi;
    if (c < 3) {
      float* a = acc + c * N + i;
      *a = accumulate ? *a + v : v;
    } else {
      phi[i] = accumulate ? phi[i] - v : -v;
    }
  }
}

// tile_flags[t]: kUniformH when every h_j of j-tile t equals the tile's
// first (bitwise), kUniformM likewise for m_j.
__global__ __launch_bounds__(256) void tile_uniform_kernel(int n_j, const float* __restrict__ h_j,
                                                           const float* __restrict__ m_j,
                                                           int j_tile, int n_tiles,
                                                           int* __restrict__ flags) {
  const int lane = threadIdx.x & (khrr::kWave - 1);
  const int t = blockIdx.x * (blockDim.x / khrr::kWave) + threadIdx.x / khrr::kWave;
  if (t >= n_tiles) return;
  const int j0 = t * j_tile;
  const int j1 = min(n_j, j0 + j_tile);
  const unsigned href = __float_as_uint(h_j[j0]);
  const unsigned mref = __float_as_uint(m_j[j0]);
  bool same_h = true, same_m = true;
  for (int j = j0 + lane; j < j1; j += khrr::kWave) {
    same_h &= __float_as_uint(h_j[j]) == href;
    same_m &= __float_as_uint(m_j[j]) == mref;
  }
  const unsigned long long bad_h = __ballot(!same_h);
  const unsigned long long bad_m = __ballot(!same_m);
  if (lane == 0)
    flags[t] = (bad_h == 0ull ? khrr::kUniformH : 0) | (bad_m == 0ull ? khrr::kUniformM : 0);
}

__global__ void fill_kernel(float* __restrict__ p, size_t n, float v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <int IPL, bool DO_ACC, bool DO_PHI, bool COMP>
void launch_grav(dim3 grid, dim3 block, hipStream_t st, int n_i, const float* pos_i,
                 const float* h_i, int n_j, const float* pos_j, const float* m_j, const float* h_j,
                 const int* flags, int groups_i, int n_split, int j_chunk, int j_tile, float* acc,
                 float* phi, float* part, int accumulate) {
  hipLaunchKernelGGL((grav_kernel<IPL, DO_ACC, DO_PHI, COMP>), grid, block, 0, st, n_i, pos_i,
                     h_i, n_j, pos_j, m_j, h_j, flags, groups_i, n_split, j_chunk, j_tile, acc,
                     phi, part, accumulate);
}

using GravLauncher = void (*)(dim3, dim3, hipStream_t, int, const float*, const float*, int,
                              const float*, const float*, const float*, const int*, int, int, int,
                              int, float*, float*, float*, int);

template <int IPL>
GravLauncher pick_mode(bool do_acc, bool do_phi, bool comp) {
  if (do_acc && do_phi)
    return comp ? launch_grav<IPL, true, true, true> : launch_grav<IPL, true, true, false>;
  if (do_acc) return comp ? launch_grav<IPL, true, false, true> : launch_grav<IPL, true, false, false>;
  return comp ? launch_grav<IPL, false, true, true> : launch_grav<IPL, false, true, false>;
}

GravLauncher pick(int ipl, bool do_acc, bool do_phi, bool comp) {
  switch (ipl) {
    case 1: return pick_mode<1>(do_acc, do_phi, comp);
    case 2: return pick_mode<2>(do_acc, do_phi, comp);
    case 4: return pick_mode<4>(do_acc, do_phi, comp);
    case 8: return pick_mode<8>(do_acc, do_phi, comp);
    default: return nullptr;
  }
}


}  // namespace

namespace khrr {

int validate_tune(khrr_ctx* ctx, const khrr_tune& t) {
  if (t.i_per_lane != 1 && t.i_per_lane != 2 && t.i_per_lane != 4 && t.i_per_lane != 8)
    return fail(ctx, KHRR_EINVAL, "tune.i_per_lane must be 1, 2, 4 or 8 (got %d)", t.i_per_lane);
  if (t.waves_per_block < 1 || t.waves_per_block > 4)
    return fail(ctx, KHRR_EINVAL, "tune.waves_per_block must be in [1,4] (got %d)",
                t.waves_per_block);
  if (t.j_split < 1 || t.j_split > 4096)
    return fail(ctx, KHRR_EINVAL, "tune.j_split must be in [1,4096] (got %d)", t.j_split);
  if (t.j_tile < 16 || t.j_tile % 16 != 0 || t.j_tile > (1 << 20))
    return fail(ctx, KHRR_EINVAL, "tune.j_tile must be a positive multiple of 16 (got %d)",
                t.j_tile);
  if (t.symmetric != 0 && t.symmetric != 1)
    return fail(ctx, KHRR_EINVAL, "tune.symmetric must be 0 or 1 (got %d)", t.symmetric);
  if (t.superblock != 0 && t.superblock != 1024 && t.superblock != 2048 && t.superblock != 4096)
    return fail(ctx, KHRR_EINVAL, "tune.superblock must be 0, 1024, 2048 or 4096 (got %d)",
                t.superblock);
  return 0;
}

// Static launch shape: the pair-symmetric kernel (8 i per lane) where it
// applies; otherwise 2 i per lane, 4 waves per block, 256-j tiles and a
// j-split that gives ~8 waves per SIMD (4 SIMDs per CU), with j-ranges of
// at least 1024 bodies, or one 256-body tile up to 8192 bodies (a small
// system's few i-groups then run many short waves instead of a few long
// ones).
khrr_tune default_tune(const khrr_ctx* ctx, int n_i, int n_j) {
  khrr_tune t;
  t.i_per_lane = 2;
  t.waves_per_block = 4;
  t.j_tile = 256;
  t.symmetric = 0;
  t.superblock = 0;
  const long long groups = (n_i + 64ll * t.i_per_lane - 1) / (64ll * t.i_per_lane);
  const long long target = 32ll * ctx->cu_count;
  long long split = groups > 0 ? (target + groups - 1) / groups : 1;
  const long long min_chunk = n_j <= 8192 ? 256 : 1024;
  const long long max_split = n_j / min_chunk > 1 ? n_j / min_chunk : 1;
  if (split > max_split) split = max_split;
  if (split < 1) split = 1;
  t.j_split = (int)split;
  return t;
}

int tune_bucket(int n) {
  int b = 0;
  while ((1ll << b) < (long long)n && b < kTuneBuckets - 1) ++b;
  return b;
}

const khrr_tune* cached_tune(const khrr_ctx* ctx, int n) {
  const int b = tune_bucket(n);
  return ctx->has_tuned[b] ? &ctx->tuned[b] : nullptr;
}

void store_tune(khrr_ctx* ctx, int n, const khrr_tune& t) {
  const int b = tune_bucket(n);
  ctx->tuned[b] = t;
  ctx->has_tuned[b] = true;
}

// Core launcher shared by the plain entry point and the autotuner.
int accel_launch(khrr_ctx* ctx, int n_i, const float* pos_i, const float* h_i, int n_j,
                 const float* pos_j, const float* m_j, const float* h_j, float* acc, float* phi,
                 unsigned flags, const khrr_tune& tune) {
  const bool do_acc = acc != nullptr, do_phi = phi != nullptr;
  const bool comp = (flags & KHRR_COMPENSATED) != 0;
  const int accumulate = (flags & KHRR_ACCUMULATE) ? 1 : 0;
  if (n_i == 0) return 0;
  if (n_j == 0) {
    if (!accumulate) {
      Timed tm(ctx, KHRR_K_GRAVITY);
      if (do_acc)
        hipLaunchKernelGGL(fill_kernel, dim3(ceil_div(3ll * n_i, 256)), dim3(256), 0, ctx->stream,
                           acc, (size_t)3 * n_i, 0.f);
      if (do_phi)
        hipLaunchKernelGGL(fill_kernel, dim3(ceil_div(n_i, 256)), dim3(256), 0, ctx->stream, phi,
                           (size_t)n_i, 0.f);
      return check_launch(ctx, "fill_kernel");
    }
    return 0;
  }
  if (tune.symmetric && n_i == n_j && pos_i == pos_j && h_i == h_j &&
      sym_eligible(n_i, tune.i_per_lane, tune.superblock)) {
    const int rc = sym_launch(ctx, n_i, pos_i, h_i, m_j, acc, phi, flags, tune.i_per_lane,
                              tune.superblock);
    if (rc != KHRR_ENOMEM) return rc;
    // no room for the partial slots: the one-sided kernel needs none
  }
  const int ipl = tune.i_per_lane;
  const int groups_i = (int)ceil_div(n_i, 64ll * ipl);
  const int j_tile = tune.j_tile;
  long long chunk = ((long long)n_j + tune.j_split - 1) / tune.j_split;
  chunk = (chunk + j_tile - 1) / j_tile * j_tile;
  const int n_split = (int)ceil_div(n_j, chunk);
  const int n_tiles = (int)ceil_div(n_j, j_tile);
  const int NC = comp ? 8 : 4;

  const size_t need = Carver::need(n_tiles, sizeof(int)) +
                      (n_split > 1 ? Carver::need((size_t)n_split * NC * n_i, sizeof(float)) : 0);
  void* base = nullptr;
  KHRR_TRY(scratch(ctx, need, &base));
  Carver cv(base);
  int* tile_flags = cv.take<int>(n_tiles);
  float* part = n_split > 1 ? cv.take<float>((size_t)n_split * NC * n_i) : nullptr;

  GravLauncher fn = pick(ipl, do_acc, do_phi, comp);
  if (!fn) return fail(ctx, KHRR_EINVAL, "unsupported i_per_lane %d", ipl);
  const int wpb = tune.waves_per_block;
  const long long waves = (long long)groups_i * n_split;
  {
    Timed tm(ctx, KHRR_K_GRAVITY);
    hipLaunchKernelGGL(tile_uniform_kernel, dim3(ceil_div(n_tiles, 4)), dim3(256), 0, ctx->stream,
                       n_j, h_j, m_j, j_tile, n_tiles, tile_flags);
    fn(dim3(ceil_div(waves, wpb)), dim3(wpb * 64), ctx->stream, n_i, pos_i, h_i, n_j, pos_j, m_j,
       h_j, tile_flags, groups_i, n_split, (int)chunk, j_tile, acc, phi, part, accumulate);
    if (n_split > 1) {
      if (comp)
        hipLaunchKernelGGL(grav_reduce<true>, dim3(ceil_div(n_i, 256)), dim3(256), 0, ctx->stream,
                           n_i, n_split, (int)do_acc, (int)do_phi, part, acc, phi, accumulate);
      else
        hipLaunchKernelGGL(grav_reduce<false>, dim3(ceil_div(n_i, 256)), dim3(256), 0,
                           ctx->stream, n_i, n_split, (int)do_acc, (int)do_phi, part, acc, phi,
                           accumulate);
    }
  }
  return check_launch(ctx, "grav_kernel");
}

}  // namespace khrr

extern "C" {

int khrr_tune_default(khrr_ctx* ctx, int n, khrr_tune* out) {
  if (!ctx || !out) return khrr::fail(ctx, KHRR_EINVAL, "null argument");
  if (n < 0) return khrr::fail(ctx, KHRR_EINVAL, "n < 0");
  *out = khrr::default_tune(ctx, n, n);
  return 0;
}

int khrr_accel(khrr_ctx* ctx, int n_i, const float* pos_i, const float* h_i, int n_j,
               const float* pos_j, const float* m_j, const float* h_j, float* acc, float* phi,
               unsigned flags, const khrr_tune* tune) {
  KHRR_TRY(khrr::enter(ctx));
  if (n_i < 0 || n_j < 0) return khrr::fail(ctx, KHRR_EINVAL, "negative size (n_i=%d n_j=%d)", n_i, n_j);
  if (flags & ~(KHRR_ACCUMULATE | KHRR_COMPENSATED))
    return khrr::fail(ctx, KHRR_EINVAL, "unknown flags 0x%x", flags);
  if (n_i == 0) return 0;  // no i-particles: nothing to write
  if (!acc && !phi) return khrr::fail(ctx, KHRR_EINVAL, "acc and phi are both NULL");
  if (n_i > 0 && (!pos_i || !h_i)) return khrr::fail(ctx, KHRR_EINVAL, "pos_i/h_i is NULL");
  if (n_j > 0 && (!pos_j || !m_j || !h_j))
    return khrr::fail(ctx, KHRR_EINVAL, "pos_j/m_j/h_j is NULL");
  khrr_tune t;
  if (tune) {
    t = *tune;
  } else if (const khrr_tune* c = khrr::cached_tune(ctx, n_j)) {
    t = *c;
  } else {
    t = khrr::default_tune(ctx, n_i, n_j);
    const bool same_set = n_i == n_j && pos_i == pos_j && h_i == h_j;
    if (same_set && khrr::sym_eligible(n_i, 8, 0) && khrr::sym_worthwhile(ctx, n_i)) {
      t.i_per_lane = 8;
      t.symmetric = 1;
    }
  }
  KHRR_TRY(khrr::validate_tune(ctx, t));
  return khrr::accel_launch(ctx, n_i, pos_i, h_i, n_j, pos_j, m_j, h_j, acc, phi, flags, t);
}

}  // extern "C"
