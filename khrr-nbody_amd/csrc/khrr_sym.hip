// khrr_sym.hip -- direct-sum gravity with each unordered pair evaluated once.
//
// Same results as ACCEL_kernel / PHI_kernel (nbody-kernel.cuh:43-142) when
// the i-set and the j-set are one system: the pair force is antisymmetric
// under the reference's symmetric softening 0.5*(eps2_i + eps2_j), so one
// rsq serves both a_i and a_j (Newton's third law).  That halves the
// transcendental, subtraction and r^2 work per directed pair.
//
// Work decomposition (gfx950):
//  * bodies are cut into superblocks of SB bodies (by default 4096 for
//    n >= 524288, 2048 from 32768, else 1024, so smaller systems still make
//    enough items; a tuned shape may pick any of the three); a workgroup
//    of SB / (64 IPL) waves takes one item (A, B) of the upper triangle,
//    A <= B;
//  * off-diagonal items: each wave owns 64*IPL i-bodies of A; the j-bodies of
//    B stream through in chunks of 64, held in LDS twice over.  At step s of
//    a chunk lane l meets j-body (l + s) mod 64, so after 64 steps every
//    lane-i has met every chunk-j once; each wave accumulates the step's
//    j-side terms onto that body's sums in its own LDS rows (the LDS loads
//    and stores issue beside the VALU work, where moving the bodies and
//    sums through the lanes with DPP cost about two FMA slots per move).
//    The W waves' j-side sums are then added in wave order;
//  * diagonal items (A, A) run the one-sided scalar-load loop;
//  * every (superblock X, body i) partial slot part[X][c][i] is written by
//    exactly one item, and a second kernel adds the slots in X order: the
//    result is deterministic and needs no atomics.
// Per unique pair with uniform mass and softening: 14 VALU ops + one
// v_rsq_f32 (plus, per IPL pairs, 3 LDS loads of the j-body and 3 loads and
// stores of its sums), i.e. about half of the one-sided cost per directed
// pair.
#include "khrr_internal.h"
#include "khrr_pair.h"

namespace {

using khrr::Accum;
using khrr::fold;
using khrr::kUniformH;
using khrr::kUniformM;


template <bool DO_ACC, bool DO_PHI, bool UNIH, bool UNIM>
__device__ __forceinline__ void pair_sym(float xj, float yj, float zj, float mj, float hj, float xi,
                                         float yi, float zi, float hi, float mi, float& tx,
                                         float& ty, float& tz, float& tp, float& bx, float& by,
                                         float& bz, float& bp) {
  const float dx = xj - xi;
  const float dy = yj - yi;
  const float dz = zj - zi;
  const float h = UNIH ? hi : hi + hj;  // UNIH: hi already holds h_i + h_j
  const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, h)));
  const float inv = __builtin_amdgcn_rsqf(r2);
  const float inv3 = inv * (inv * inv);
  if (DO_ACC) {
    const float si = UNIM ? inv3 : mj * inv3;
    const float sj = UNIM ? inv3 : mi * inv3;
    tx = __builtin_fmaf(dx, si, tx);
    ty = __builtin_fmaf(dy, si, ty);
    tz = __builtin_fmaf(dz, si, tz);
    bx = __builtin_fmaf(dx, sj, bx);
    by = __builtin_fmaf(dy, sj, by);
    bz = __builtin_fmaf(dz, sj, bz);
  }
  if (DO_PHI) {
    tp += UNIM ? inv : mj * inv;
    bp += UNIM ? inv : mi * inv;
  }
}

// Number of j-side sums per body: ax, ay, az and/or the potential, which
// takes slot 3 next to the acceleration and slot 0 alone.
template <bool DO_ACC, bool DO_PHI>
constexpr int kSums = (DO_ACC ? 3 : 0) + (DO_PHI ? 1 : 0);
template <bool DO_ACC>
constexpr int kPhiSlot = DO_ACC ? 3 : 0;

// One 64-body j chunk in LDS, twice over: entries e and e + 64 both hold
// body e, so step s reads entry lane + s without wrapping.
struct JChunk {
  float x[2 * khrr::kWave], y[2 * khrr::kWave], z[2 * khrr::kWave];
  float m[2 * khrr::kWave], h[2 * khrr::kWave];
};

// LDS floats of the off-diagonal loop's j-side sums: per wave two sets (one
// per chunk parity) of kSums rows of 128 entries, in which entries e and
// e + 64 together hold body e's sums.
template <int W, bool DO_ACC, bool DO_PHI>
constexpr int kJSumFloats = W * 2 * kSums<DO_ACC, DO_PHI> * 2 * khrr::kWave;

// 64 steps over one chunk of j-bodies in LDS: at step s lane l pairs its
// i-bodies with j-body (l + s) mod 64, accumulating onto that body's entry
// l + s (loaded before the step's pairs, stored after them).  Lane l + 1
// wrote the entry the step before, so a compiler barrier after each step
// keeps the loads and stores of the sums in program order, in which the LDS
// applies one wave's operations.
template <int IPL, bool DO_ACC, bool DO_PHI, bool UNIH, bool UNIM>
__device__ __forceinline__ void chunk64_lds(const JChunk& J, float* sums, int lane,
                                            const float (&xi)[IPL], const float (&yi)[IPL],
                                            const float (&zi)[IPL], const float (&hi)[IPL],
                                            const float (&mi)[IPL], float (&tx)[IPL],
                                            float (&ty)[IPL], float (&tz)[IPL], float (&tp)[IPL]) {
  constexpr int R = 2 * khrr::kWave;
  const float* px = J.x + lane;
  const float* py = J.y + lane;
  const float* pz = J.z + lane;
  const float* pm = J.m + lane;
  const float* ph = J.h + lane;
  float* sx = sums + 0 * R + lane;
  float* sy = sums + 1 * R + lane;
  float* sz = sums + 2 * R + lane;
  float* sp = sums + kPhiSlot<DO_ACC> * R + lane;
#pragma unroll 4
  for (int s = 0; s < khrr::kWave; ++s) {
    const float xj = px[s], yj = py[s], zj = pz[s];
    const float mj = UNIM ? 0.f : pm[s];
    const float hj = UNIH ? 0.f : ph[s];
    // the step's j-side terms go straight onto the body's running sums
    float bx = DO_ACC ? sx[s] : 0.f, by = DO_ACC ? sy[s] : 0.f, bz = DO_ACC ? sz[s] : 0.f;
    float bp = DO_PHI ? sp[s] : 0.f;
#pragma unroll
    for (int k = 0; k < IPL; ++k)
      pair_sym<DO_ACC, DO_PHI, UNIH, UNIM>(xj, yj, zj, mj, hj, xi[k], yi[k], zi[k], hi[k], mi[k],
                                           tx[k], ty[k], tz[k], tp[k], bx, by, bz, bp);
    if (DO_ACC) {
      sx[s] = bx;
      sy[s] = by;
      sz[s] = bz;
    }
    if (DO_PHI) sp[s] = bp;
    asm volatile("" ::: "memory");
  }
}

// One (A, B) item of one system of n bodies whose planes start at xs/ys/zs,
// m, h (any stride between planes); sb_flags[X] are the system's superblock
// flags and part its slot block, laid out part[(X*4 + c)*n + i] with
// c = ax, ay, az (acceleration from the bodies of superblock X) and
// 3 = sum_j m_j / r (positive).
template <int IPL, int W, bool DO_ACC, bool DO_PHI, bool COMP>
__device__ __forceinline__ void sym_block(
    int n, int A, int B, const float* __restrict__ xs, const float* __restrict__ ys,
    const float* __restrict__ zs, const float* __restrict__ m, const float* __restrict__ h,
    const int* __restrict__ sb_flags, float* __restrict__ part, JChunk (&jc)[2],
    float* __restrict__ jsum, float* __restrict__ lo_all) {
  static_assert(W >= 4, "the j-side sum uses 4 x 64 threads");
  constexpr int SB = khrr::kWave * IPL * W;  // bodies per superblock
  const int lane = threadIdx.x & (khrr::kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / khrr::kWave);
  const size_t N = (size_t)n;
  const int flags = sb_flags[A] & sb_flags[B];

  float xi[IPL], yi[IPL], zi[IPL], hi[IPL], mi[IPL];
  int idx[IPL];
#pragma unroll
  for (int k = 0; k < IPL; ++k) {
    idx[k] = A * SB + wave * (khrr::kWave * IPL) + k * khrr::kWave + lane;
    const bool ok = idx[k] < n;
    // an absent i sits at 1e30: r^2 overflows to inf, rsq gives 0 and every
    // term it adds to a j-body is exactly 0
    xi[k] = ok ? xs[idx[k]] : 1e30f;
    yi[k] = ok ? ys[idx[k]] : 1e30f;
    zi[k] = ok ? zs[idx[k]] : 1e30f;
    hi[k] = ok ? h[idx[k]] : 1.f;
    mi[k] = ok ? m[idx[k]] : 0.f;
  }
  Accum<IPL> acc;
  acc.zero();
  // compensated form: the TwoSum error terms live in LDS (this lane's slot
  // of a [kSums][IPL][64] block per wave), so the registers hold the same
  // working set as the plain form
  constexpr int NS = kSums<DO_ACC, DO_PHI>;
  constexpr int PS = kPhiSlot<DO_ACC>;
  float* lo = COMP ? lo_all + (size_t)wave * NS * IPL * khrr::kWave + lane : nullptr;
  if (COMP) {
#pragma unroll
    for (int q = 0; q < NS * IPL; ++q) lo[q * khrr::kWave] = 0.f;
  }

  if (A == B) {
    // diagonal: one-sided over the superblock's own bodies (j = i included)
    const int j0 = A * SB, j1 = min(n, j0 + SB);
    for (int jt = j0; jt < j1; jt += 256)
      khrr::tile_step<IPL, DO_ACC, DO_PHI, COMP>(jt, min(j1, jt + 256), sb_flags[A], xs, ys, zs,
                                                 m, h, xi, yi, zi, hi, acc, lo);
  } else {
    // Only the specialised loop (softening and mass both uniform over A and
    // B) leaves the masses out of the pair terms; the generic loop applies
    // m_j and m_i per pair, so its sums take no scale afterwards.
    const bool uni = (flags & kUniformH) != 0 && (flags & kUniformM) != 0;
    const int j0 = B * SB, j1 = min(n, j0 + SB);
    const float hsum = h[A * SB] + h[j0];
    const float mA = m[A * SB], mB = m[j0];
    float hh[IPL];  // h_i + h_j of the uniform path
#pragma unroll
    for (int k = 0; k < IPL; ++k) hh[k] = hsum;
    // The chunk's bodies, or the absent-body sentinel past j1 (an absent j
    // sits at 1e30: every term it gets or gives is exactly 0).
    auto load_j = [&](int j, float& x, float& y, float& z, float& mm, float& hh_) {
      const bool vj = j < j1;
      x = vj ? xs[j] : 1e30f;
      y = vj ? ys[j] : 1e30f;
      z = vj ? zs[j] : 1e30f;
      mm = vj ? m[j] : 0.f;
      hh_ = vj ? h[j] : 1.f;
    };
    auto store_j = [&](JChunk& C, float x, float y, float z, float mm, float hh_) {
      C.x[lane] = C.x[lane + khrr::kWave] = x;
      C.y[lane] = C.y[lane + khrr::kWave] = y;
      C.z[lane] = C.z[lane + khrr::kWave] = z;
      C.m[lane] = C.m[lane + khrr::kWave] = mm;
      C.h[lane] = C.h[lane + khrr::kWave] = hh_;
    };
    // Wave 0 writes chunk k + 1 into buffer (k + 1) & 1 before its steps of
    // chunk k; the barrier of chunk k publishes it.  The waves last read
    // that buffer in chunk k - 1, before the barrier wave 0 passed first.
    // (Wave 0 waits for those loads; the other workgroups' waves on its SIMD
    // issue meanwhile.)
    if (wave == 0) {
      float x, y, z, mm, hh_;
      load_j(j0 + lane, x, y, z, mm, hh_);
      store_j(jc[0], x, y, z, mm, hh_);
    }
    __syncthreads();
    for (int c0 = j0; c0 < j1; c0 += khrr::kWave) {
      const int par = ((c0 - j0) / khrr::kWave) & 1;
      const bool next = c0 + khrr::kWave < j1;  // uniform
      if (wave == 0 && next) {
        float x, y, z, mm, hh_;
        load_j(c0 + khrr::kWave + lane, x, y, z, mm, hh_);
        store_j(jc[par ^ 1], x, y, z, mm, hh_);
      }
      // this wave's j-side sums of parity par, cleared: their readers of two
      // chunks ago finished before the last barrier
      constexpr int R = 2 * khrr::kWave;  // entries per row of sums
      float* S = jsum + (size_t)(wave * 2 + par) * NS * R;
#pragma unroll
      for (int q = 0; q < NS; ++q) S[q * R + lane] = S[q * R + lane + khrr::kWave] = 0.f;
      // other lanes of this wave read and add to these entries next
      asm volatile("" ::: "memory");
      float tx[IPL], ty[IPL], tz[IPL], tp[IPL];
#pragma unroll
      for (int k = 0; k < IPL; ++k) tx[k] = ty[k] = tz[k] = tp[k] = 0.f;
      if (uni)
        chunk64_lds<IPL, DO_ACC, DO_PHI, true, true>(jc[par], S, lane, xi, yi, zi, hh, mi, tx, ty,
                                                     tz, tp);
      else
        chunk64_lds<IPL, DO_ACC, DO_PHI, false, false>(jc[par], S, lane, xi, yi, zi, hi, mi, tx,
                                                       ty, tz, tp);
      const float si = uni ? mB : 1.f;
#pragma unroll
      for (int k = 0; k < IPL; ++k) {
        if (COMP) {
          if (DO_ACC) {
            khrr::fold_lds(acc.ax[k], lo + (0 * IPL + k) * khrr::kWave, tx[k] * si);
            khrr::fold_lds(acc.ay[k], lo + (1 * IPL + k) * khrr::kWave, ty[k] * si);
            khrr::fold_lds(acc.az[k], lo + (2 * IPL + k) * khrr::kWave, tz[k] * si);
          }
          if (DO_PHI) khrr::fold_lds(acc.ap[k], lo + (PS * IPL + k) * khrr::kWave, tp[k] * si);
          continue;
        }
        if (DO_ACC) {
          fold(acc.ax[k], acc.lx[k], tx[k] * si, COMP);
          fold(acc.ay[k], acc.ly[k], ty[k] * si, COMP);
          fold(acc.az[k], acc.lz[k], tz[k] * si, COMP);
        }
        if (DO_PHI) fold(acc.ap[k], acc.lp[k], tp[k] * si, COMP);
      }
      // j-side: add the W waves' sums in wave order.  Chunk k adds into
      // set k & 1; its readers finish before the barrier of chunk k + 1,
      // which every writer of chunk k + 2 passes first, so one barrier per
      // chunk suffices.
      __syncthreads();
      if (threadIdx.x < 4 * khrr::kWave) {
        const int c = threadIdx.x / khrr::kWave;
        const int l = threadIdx.x % khrr::kWave;
        const bool want = c < 3 ? DO_ACC : DO_PHI;
        if (want && c0 + l < j1) {
          const int q = c < 3 ? c : PS;
          float s = 0.f;
          for (int w = 0; w < W; ++w) {
            const float* Sw = jsum + ((size_t)(w * 2 + par) * NS + q) * R;
            s += Sw[l] + Sw[l + khrr::kWave];
          }
          const float sj = uni ? mA : 1.f;
          // a_j gets -(sum m_i d_ij / r^3): d_ij points from i to j
          part[((size_t)A * 4 + c) * N + c0 + l] = c < 3 ? -(s * sj) : s * sj;
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < IPL; ++k) {
    if (idx[k] >= n) continue;
    const size_t base = (size_t)B * 4 * N + idx[k];
    if (DO_ACC) {
      part[base] = COMP ? acc.ax[k] + lo[(0 * IPL + k) * khrr::kWave] : acc.ax[k];
      part[base + N] = COMP ? acc.ay[k] + lo[(1 * IPL + k) * khrr::kWave] : acc.ay[k];
      part[base + 2 * N] = COMP ? acc.az[k] + lo[(2 * IPL + k) * khrr::kWave] : acc.az[k];
    }
    if (DO_PHI)
      part[base + 3 * N] = COMP ? acc.ap[k] + lo[(PS * IPL + k) * khrr::kWave] : acc.ap[k];
  }
}

template <int IPL, int W, bool DO_ACC, bool DO_PHI, bool COMP>
__global__ __launch_bounds__(khrr::kWave* W) void sym_kernel(int n, int n_sb,
                                                             const float* __restrict__ pos,
                                                             const float* __restrict__ m,
                                                             const float* __restrict__ h,
                                                             const int* __restrict__ sb_flags,
                                                             float* __restrict__ part) {
  __shared__ JChunk jc[2];
  __shared__ float jsum[kJSumFloats<W, DO_ACC, DO_PHI>];
  __shared__ float lo_all[COMP ? W * kSums<DO_ACC, DO_PHI> * IPL * khrr::kWave : 1];
  // item -> (A, B), row-major over the upper triangle
  int t = blockIdx.x, A = 0, row = n_sb;
  while (t >= row) {
    t -= row;
    ++A;
    --row;
  }
  const size_t N = (size_t)n;
  sym_block<IPL, W, DO_ACC, DO_PHI, COMP>(n, A, A + t, pos, pos + N, pos + 2 * N, m, h, sb_flags,
                                          part, jc, jsum, lo_all);
}

// Batched form: workgroup b takes items[b], an (A, B) item of one system of
// a segment table (planes of stride ld); every item of one launch has the
// superblock size 64 * IPL * W.
template <int IPL, int W, bool DO_ACC, bool DO_PHI, bool COMP>
__global__ __launch_bounds__(khrr::kWave* W) void sym_items_kernel(
    const khrr::SymItem* __restrict__ items, int ld, const float* __restrict__ pos,
    const float* __restrict__ m, const float* __restrict__ h, const int* __restrict__ flags,
    float* __restrict__ part) {
  __shared__ JChunk jc[2];
  __shared__ float jsum[kJSumFloats<W, DO_ACC, DO_PHI>];
  __shared__ float lo_all[COMP ? W * kSums<DO_ACC, DO_PHI> * IPL * khrr::kWave : 1];
  const khrr::SymItem it = items[blockIdx.x];
  const size_t b = (size_t)it.base;
  sym_block<IPL, W, DO_ACC, DO_PHI, COMP>(it.n, it.A, it.B, pos + b, pos + ld + b,
                                          pos + 2 * (size_t)ld + b, m + b, h + b,
                                          flags + it.flag_base, part + it.part, jc, jsum, lo_all);
}

// acc[i] (+)= sum_X part[X][0..2][i];  phi[i] (-)= sum_X part[X][3][i]
// Folds slots p[0], p[stride], ... p[(count-1)*stride] in slot order, with
// eight loads in flight before their adds (the order of the adds, and so
// the result, is that of a one-at-a-time walk).
template <bool COMP>
__device__ __forceinline__ void fold_slots(const float* __restrict__ p, size_t stride, int count,
                                           float& hi, float& lo) {
  constexpr int kU = 8;
  int X = 0;
  for (; X + kU <= count; X += kU) {
    float t[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) t[u] = p[(size_t)(X + u) * stride];
#pragma unroll
    for (int u = 0; u < kU; ++u) fold(hi, lo, t[u], COMP);
  }
  for (; X < count; ++X) fold(hi, lo, p[(size_t)X * stride], COMP);
}

template <bool COMP>
__global__ __launch_bounds__(256) void sym_reduce(int n, int n_sb, int do_acc, int do_phi,
                                                  const float* __restrict__ part,
                                                  float* __restrict__ acc,
                                                  float* __restrict__ phi, int accumulate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t N = (size_t)n;
  for (int c = 0; c < 4; ++c) {
    if ((c < 3 && !do_acc) || (c == 3 && !do_phi)) continue;
    float hi = 0.f, lo = 0.f;
    fold_slots<COMP>(part + (size_t)c * N + i, 4 * N, n_sb, hi, lo);
    const float v = COMP ? hi + lo : hi;
    if (c < 3) {
      float* a = acc + c * N + i;
      *a = accumulate ? *a + v : v;
    } else {
      phi[i] = accumulate ? phi[i] - v : -v;
    }
  }
}

// Batched form of sym_reduce: blockIdx.y selects the system.
template <bool COMP>
__global__ __launch_bounds__(256) void sym_reduce_sys(const khrr::SymSys* __restrict__ sys, int ld,
                                                      int do_acc, int do_phi,
                                                      const float* __restrict__ part_all,
                                                      float* __restrict__ acc,
                                                      float* __restrict__ phi, int accumulate) {
  const khrr::SymSys S = sys[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S.n) return;
  const size_t N = (size_t)S.n;
  const float* part = part_all + S.part;
  for (int c = 0; c < 4; ++c) {
    if ((c < 3 && !do_acc) || (c == 3 && !do_phi)) continue;
    float hi = 0.f, lo = 0.f;
    fold_slots<COMP>(part + (size_t)c * N + i, 4 * N, S.n_sb, hi, lo);
    const float v = COMP ? hi + lo : hi;
    const size_t g = (size_t)S.base + i;
    if (c < 3) {
      float* a = acc + (size_t)c * ld + g;
      *a = accumulate ? *a + v : v;
    } else {
      phi[g] = accumulate ? phi[g] - v : -v;
    }
  }
}

// sb_flags[X]: kUniformH / kUniformM over superblock X.
__global__ __launch_bounds__(256) void sb_flags_kernel(int n, int n_sb, int sb,
                                                       const float* __restrict__ h,
                                                       const float* __restrict__ m,
                                                       int* __restrict__ flags) {
  const int lane = threadIdx.x & (khrr::kWave - 1);
  const int X = blockIdx.x * (blockDim.x / khrr::kWave) + threadIdx.x / khrr::kWave;
  if (X >= n_sb) return;
  const int b = X * sb, e = min(n, b + sb);
  const unsigned href = __float_as_uint(h[b]), mref = __float_as_uint(m[b]);
  bool sh = true, sm = true;
  for (int j = b + lane; j < e; j += khrr::kWave) {
    sh &= __float_as_uint(h[j]) == href;
    sm &= __float_as_uint(m[j]) == mref;
  }
  const unsigned long long bh = __ballot(!sh), bm = __ballot(!sm);
  if (lane == 0) flags[X] = (bh == 0ull ? kUniformH : 0) | (bm == 0ull ? kUniformM : 0);
}

template <int IPL, int W, bool A, bool P, bool C>
void launch_sym(hipStream_t st, int n, int n_sb, const float* pos, const float* m, const float* h,
                const int* flags, float* part) {
  const unsigned items = (unsigned)((long long)n_sb * (n_sb + 1) / 2);
  hipLaunchKernelGGL((sym_kernel<IPL, W, A, P, C>), dim3(items), dim3(khrr::kWave * W), 0, st, n,
                     n_sb, pos, m, h, flags, part);
}

using SymLauncher = void (*)(hipStream_t, int, int, const float*, const float*, const float*,
                             const int*, float*);

template <int IPL, int W>
SymLauncher pick_sym_mode(bool a, bool p, bool c) {
  if (a && p) return c ? launch_sym<IPL, W, true, true, true> : launch_sym<IPL, W, true, true, false>;
  if (a) return c ? launch_sym<IPL, W, true, false, true> : launch_sym<IPL, W, true, false, false>;
  return c ? launch_sym<IPL, W, false, true, true> : launch_sym<IPL, W, false, true, false>;
}

// ipl 8 with 4096- or 2048-body superblocks, ipl 4 with 4096, 2048 or 1024
// (1024-body superblocks always run with ipl 4: khrr::sym_ipl).
SymLauncher pick_sym(int ipl, int sb, bool a, bool p, bool c) {
  if (sb == 1024) return pick_sym_mode<4, 4>(a, p, c);
  if (ipl == 8) return sb == 4096 ? pick_sym_mode<8, 8>(a, p, c) : pick_sym_mode<8, 4>(a, p, c);
  return sb == 4096 ? pick_sym_mode<4, 16>(a, p, c) : pick_sym_mode<4, 8>(a, p, c);
}

template <int IPL, int W, bool A, bool P, bool C>
void launch_sym_items(hipStream_t st, int n_items, const khrr::SymItem* items, int ld,
                      const float* pos, const float* m, const float* h, const int* flags,
                      float* part) {
  hipLaunchKernelGGL((sym_items_kernel<IPL, W, A, P, C>), dim3(n_items), dim3(khrr::kWave * W), 0,
                     st, items, ld, pos, m, h, flags, part);
}

using SymItemsLauncher = void (*)(hipStream_t, int, const khrr::SymItem*, int, const float*,
                                  const float*, const float*, const int*, float*);

template <int IPL, int W>
SymItemsLauncher pick_sym_items_mode(bool a, bool p, bool c) {
  if (a && p)
    return c ? launch_sym_items<IPL, W, true, true, true> : launch_sym_items<IPL, W, true, true, false>;
  if (a)
    return c ? launch_sym_items<IPL, W, true, false, true> : launch_sym_items<IPL, W, true, false, false>;
  return c ? launch_sym_items<IPL, W, false, true, true> : launch_sym_items<IPL, W, false, true, false>;
}

}  // namespace

namespace khrr {

// Partial-slot buffer cap: 16 B per body per superblock is 2.4 GB at
// 786432 bodies and 38.6 GB at 3145728 (of 288 GB HBM).
constexpr size_t kSymPartBudget = size_t(40) << 30;

// 2048-body superblocks give four times the items of 4096 (and four
// 4-wave workgroups per CU instead of two 8-wave ones), so a lone system
// below a few 10^5 bodies leaves a smaller share of the launch to its last,
// partly filled wave of workgroups; from 524288 bodies on the 4096 form has
// items enough and half the partial-slot memory.  Below 32768 bodies,
// 1024-body superblocks (4 i per lane, 4 waves) keep 16 or more of them.
// A tuned launch shape may ask for any of the three sizes instead.
int sym_superblock(int n, int requested) {
  if (requested == 1024 || requested == 2048 || requested == 4096) return requested;
  return n >= 524288 ? 4096 : n >= 32768 ? 2048 : 1024;
}

int sym_ipl(int sb, int ipl) { return sb == 1024 ? 4 : ipl; }

bool sym_eligible(int n, int ipl, int requested) {
  if (ipl != 4 && ipl != 8) return false;
  if (requested != 0 && requested != 1024 && requested != 2048 && requested != 4096) return false;
  const int sb = sym_superblock(n, requested);
  const long long n_sb = (n + (long long)sb - 1) / sb;
  if (n_sb < 2) return false;
  return (size_t)n_sb * 4 * (size_t)n * sizeof(float) <= kSymPartBudget;
}

// Enough (A, B) items to fill the device (about two workgroups per CU).
bool sym_worthwhile(const khrr_ctx* ctx, int n) {
  const int sb = sym_superblock(n, 0);
  const long long n_sb = (n + (long long)sb - 1) / sb;
  return n_sb * (n_sb + 1) / 2 >= 2ll * ctx->cu_count;
}

int sym_launch(khrr_ctx* ctx, int n, const float* pos, const float* h, const float* m, float* acc,
               float* phi, unsigned flags, int ipl, int requested) {
  const bool do_acc = acc != nullptr, do_phi = phi != nullptr;
  const bool comp = (flags & KHRR_COMPENSATED) != 0;
  const int accumulate = (flags & KHRR_ACCUMULATE) ? 1 : 0;
  const int sb = sym_superblock(n, requested);
  const int n_sb = (n + sb - 1) / sb;
  ipl = sym_ipl(sb, ipl);
  const size_t part_elems = (size_t)n_sb * 4 * n;
  void* base = nullptr;
  KHRR_TRY(scratch(ctx, Carver::need(n_sb, sizeof(int)) + Carver::need(part_elems, sizeof(float)),
                   &base));
  Carver cv(base);
  int* sbf = cv.take<int>(n_sb);
  float* part = cv.take<float>(part_elems);
  SymLauncher fn = pick_sym(ipl, sb, do_acc, do_phi, comp);
  Timed tm(ctx, KHRR_K_GRAVITY);
  hipLaunchKernelGGL(sb_flags_kernel, dim3(ceil_div(n_sb, 4)), dim3(256), 0, ctx->stream, n, n_sb,
                     sb, h, m, sbf);
  fn(ctx->stream, n, n_sb, pos, m, h, sbf, part);
  if (comp)
    hipLaunchKernelGGL(sym_reduce<true>, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, n,
                       n_sb, (int)do_acc, (int)do_phi, part, acc, phi, accumulate);
  else
    hipLaunchKernelGGL(sym_reduce<false>, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, n,
                       n_sb, (int)do_acc, (int)do_phi, part, acc, phi, accumulate);
  return check_launch(ctx, "sym_kernel");
}

int sym_items_launch(khrr_ctx* ctx, int n_big, int n_mid, int n_small, const SymItem* items,
                     int n_sys, const SymSys* sys, int max_n, int ld, const float* pos,
                     const float* h, const float* m, const int* sb_flags, float* part, float* acc,
                     float* phi, unsigned flags) {
  if (n_big + n_mid + n_small == 0) return 0;
  const bool do_acc = acc != nullptr, do_phi = phi != nullptr;
  const bool comp = (flags & KHRR_COMPENSATED) != 0;
  const int accumulate = (flags & KHRR_ACCUMULATE) ? 1 : 0;
  if (n_big > 0)  // 4096-body superblocks
    pick_sym_items_mode<8, 8>(do_acc, do_phi, comp)(ctx->stream, n_big, items, ld, pos, m, h,
                                                    sb_flags, part);
  if (n_mid > 0)  // 2048
    pick_sym_items_mode<8, 4>(do_acc, do_phi, comp)(ctx->stream, n_mid, items + n_big, ld, pos, m,
                                                    h, sb_flags, part);
  if (n_small > 0)  // 1024
    pick_sym_items_mode<4, 4>(do_acc, do_phi, comp)(ctx->stream, n_small, items + n_big + n_mid, ld,
                                                    pos, m, h, sb_flags, part);
  const dim3 grid(ceil_div(max_n, 256), n_sys);
  if (comp)
    hipLaunchKernelGGL(sym_reduce_sys<true>, grid, dim3(256), 0, ctx->stream, sys, ld, (int)do_acc,
                       (int)do_phi, part, acc, phi, accumulate);
  else
    hipLaunchKernelGGL(sym_reduce_sys<false>, grid, dim3(256), 0, ctx->stream, sys, ld, (int)do_acc,
                       (int)do_phi, part, acc, phi, accumulate);
  return check_launch(ctx, "sym_items_kernel");
}

}  // namespace khrr
