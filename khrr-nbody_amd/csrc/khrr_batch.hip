// khrr_batch.hip -- many independent systems of uneven N in one persistent
// launch, and the project's launch-shape autotuner.
//
// The reference runs one system per process (nbody.cu:1560 launches
// ACCEL_kernel with N/512 blocks), which leaves most of 256 CUs idle for
// N below a few 10^5.  Here the host turns the segment table into a work
// list of (system, i-group) items, longest system first (LPT order), and
// one persistent grid of waves drains it through an atomic counter.  Each
// item runs the same pair loop as the plain kernel over its own system's
// j-range, so every system's result is independent of what else is in the
// batch (no shared reductions), bit for bit.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "khrr_internal.h"
#include "khrr_pair.h"

namespace {

using khrr::Accum;

template <int IPL, bool DO_ACC, bool DO_PHI, bool COMP>
__global__ __launch_bounds__(256) void batched_kernel(
    int n_items, const int2* __restrict__ items, const int* __restrict__ sys_off,
    const int* __restrict__ sys_flags, int n_total, const float* __restrict__ pos,
    const float* __restrict__ hh_all, const float* __restrict__ m, int j_tile,
    int* __restrict__ counter, float* __restrict__ acc, float* __restrict__ phi,
    int accumulate) {
  const int lane = threadIdx.x & (khrr::kWave - 1);
  const float* xa = pos;
  const float* ya = pos + n_total;
  const float* za = pos + 2 * (size_t)n_total;
  for (;;) {
    int got = 0;
    if (lane == 0) got = atomicAdd(counter, 1);
    const int item = __builtin_amdgcn_readfirstlane(got);
    if (item >= n_items) break;
    const int2 it = items[item];
    const int base = sys_off[it.x];
    const int n = sys_off[it.x + 1] - base;
    const int flags = sys_flags[it.x];

    float xi[IPL], yi[IPL], zi[IPL], hi[IPL];
    int idx[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
      const int loc = it.y * (IPL * khrr::kWave) + k * khrr::kWave + lane;
      idx[k] = loc < n ? base + loc : -1;
      const bool ok = idx[k] >= 0;
      xi[k] = ok ? xa[idx[k]] : 0.f;
      yi[k] = ok ? ya[idx[k]] : 0.f;
      zi[k] = ok ? za[idx[k]] : 0.f;
      hi[k] = ok ? hh_all[idx[k]] : 1.f;
    }
    Accum<IPL> a;
    a.zero();
    const int j1 = base + n;
    for (int jt = base; jt < j1; jt += j_tile) {
      const int je = min(j1, jt + j_tile);
      khrr::tile_step<IPL, DO_ACC, DO_PHI, COMP>(jt, je, flags, xa, ya, za, m, hh_all, xi, yi, zi,
                                                 hi, a);
    }
#pragma unroll
    for (int k = 0; k < IPL; ++k)
      if (idx[k] >= 0)
        khrr::store_final<IPL, DO_ACC, DO_PHI, COMP>(a, k, idx[k], (size_t)n_total, acc, phi,
                                                     accumulate);
  }
}

// sys_flags[s]: kUniformH / kUniformM when all h / all m of system s are
// bitwise equal.  A system's flags depend on its own bodies only, so its
// result does not depend on the rest of the batch.
__global__ __launch_bounds__(256) void sys_uniform_kernel(int n_sys, const int* __restrict__ off,
                                                          const float* __restrict__ h,
                                                          const float* __restrict__ m,
                                                          int* __restrict__ flags) {
  const int lane = threadIdx.x & (khrr::kWave - 1);
  const int s = blockIdx.x * (blockDim.x / khrr::kWave) + threadIdx.x / khrr::kWave;
  if (s >= n_sys) return;
  const int b = off[s], e = off[s + 1];
  bool same_h = true, same_m = true;
  if (e > b) {
    const unsigned href = __float_as_uint(h[b]);
    const unsigned mref = __float_as_uint(m[b]);
    for (int j = b + lane; j < e; j += khrr::kWave) {
      same_h &= __float_as_uint(h[j]) == href;
      same_m &= __float_as_uint(m[j]) == mref;
    }
  }
  const unsigned long long bad_h = __ballot(!same_h);
  const unsigned long long bad_m = __ballot(!same_m);
  if (lane == 0)
    flags[s] = (bad_h == 0ull ? khrr::kUniformH : 0) | (bad_m == 0ull ? khrr::kUniformM : 0);
}

__global__ void set_word(int* __restrict__ p, int v) { *p = v; }

using BatchLauncher = void (*)(dim3, dim3, hipStream_t, int, const int2*, const int*, const int*,
                               int, const float*, const float*, const float*, int, int*, float*,
                               float*, int);

template <int IPL, bool A, bool P, bool C>
void launch_batched(dim3 g, dim3 b, hipStream_t st, int n_items, const int2* items,
                    const int* off, const int* uni, int n_total, const float* pos, const float* h,
                    const float* m, int j_tile, int* counter, float* acc, float* phi, int accum) {
  hipLaunchKernelGGL((batched_kernel<IPL, A, P, C>), g, b, 0, st, n_items, items, off, uni,
                     n_total, pos, h, m, j_tile, counter, acc, phi, accum);
}

template <int IPL>
BatchLauncher pick_batched_mode(bool a, bool p, bool c) {
  if (a && p) return c ? launch_batched<IPL, true, true, true> : launch_batched<IPL, true, true, false>;
  if (a) return c ? launch_batched<IPL, true, false, true> : launch_batched<IPL, true, false, false>;
  return c ? launch_batched<IPL, false, true, true> : launch_batched<IPL, false, true, false>;
}

BatchLauncher pick_batched(int ipl, bool a, bool p, bool c) {
  switch (ipl) {
    case 1: return pick_batched_mode<1>(a, p, c);
    case 2: return pick_batched_mode<2>(a, p, c);
    case 4: return pick_batched_mode<4>(a, p, c);
    case 8: return pick_batched_mode<8>(a, p, c);
    default: return nullptr;
  }
}

// Builds (or reuses) the device work list for this composition.
int prepare_items(khrr_ctx* ctx, int n_sys, const int* host_off, const int* dev_off,
                  const float* h, const float* m, int ipl, int2** items, int** uni,
                  int* n_items) {
  const bool same = ctx->batch_ipl == ipl && (int)ctx->batch_off.size() == n_sys + 1 &&
                    std::equal(ctx->batch_off.begin(), ctx->batch_off.end(), host_off);
  if (!same) {
    std::vector<int> order(n_sys);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return host_off[a + 1] - host_off[a] > host_off[b + 1] - host_off[b];
    });
    std::vector<int2> list;
    for (int s : order) {
      const int n = host_off[s + 1] - host_off[s];
      const int groups = (int)khrr::ceil_div(n, 64ll * ipl);
      for (int g = 0; g < groups; ++g) list.push_back(make_int2(s, g));
    }
    const size_t item_bytes = khrr::Carver::need(list.size() > 0 ? list.size() : 1, sizeof(int2));
    const size_t need = item_bytes + khrr::Carver::need(n_sys > 0 ? n_sys : 1, sizeof(int));
    KHRR_HIP(ctx, hipStreamSynchronize(ctx->stream));  // old list/pinned copy may be in flight
    if (need > ctx->batch_dev_bytes) {
      if (ctx->batch_dev) KHRR_HIP(ctx, hipFree(ctx->batch_dev));
      ctx->batch_dev = nullptr;
      ctx->batch_dev_bytes = 0;
      KHRR_HIP(ctx, hipMalloc(&ctx->batch_dev, need));
      ctx->batch_dev_bytes = need;
    }
    if (list.size() * sizeof(int2) > ctx->batch_pinned_bytes) {
      if (ctx->batch_pinned) KHRR_HIP(ctx, hipHostFree(ctx->batch_pinned));
      ctx->batch_pinned = nullptr;
      ctx->batch_pinned_bytes = 0;
      KHRR_HIP(ctx, hipHostMalloc(&ctx->batch_pinned, list.size() * sizeof(int2)));
      ctx->batch_pinned_bytes = list.size() * sizeof(int2);
    }
    if (!list.empty()) {
      std::memcpy(ctx->batch_pinned, list.data(), list.size() * sizeof(int2));
      KHRR_HIP(ctx, hipMemcpyAsync(ctx->batch_dev, ctx->batch_pinned, list.size() * sizeof(int2),
                                   hipMemcpyHostToDevice, ctx->stream));
    }
    ctx->batch_off.assign(host_off, host_off + n_sys + 1);
    ctx->batch_ipl = ipl;
    ctx->batch_items = (int)list.size();
  }
  *items = static_cast<int2*>(ctx->batch_dev);
  *uni = reinterpret_cast<int*>(static_cast<char*>(ctx->batch_dev) +
                                khrr::Carver::need(ctx->batch_items > 0 ? ctx->batch_items : 1,
                                                   sizeof(int2)));
  *n_items = ctx->batch_items;
  // Softening and masses can change between calls with the same composition.
  if (n_sys > 0)
    hipLaunchKernelGGL(sys_uniform_kernel, dim3(khrr::ceil_div(n_sys, 4)), dim3(256), 0,
                       ctx->stream, n_sys, dev_off, h, m, *uni);
  return khrr::check_launch(ctx, "sys_uniform_kernel");
}

int batched_launch(khrr_ctx* ctx, int n_sys, const int* sys_off, const int* host_sys_off,
                   const float* pos, const float* h, const float* m, float* acc, float* phi,
                   unsigned flags, const khrr_tune& t) {
  const int n_total = host_sys_off[n_sys];
  if (n_total == 0 || n_sys == 0) return 0;
  int2* items = nullptr;
  int* uni = nullptr;
  int n_items = 0;
  KHRR_TRY(prepare_items(ctx, n_sys, host_sys_off, sys_off, h, m, t.i_per_lane, &items, &uni,
                         &n_items));
  BatchLauncher fn = pick_batched(t.i_per_lane, acc != nullptr, phi != nullptr,
                                  (flags & KHRR_COMPENSATED) != 0);
  if (!fn) return khrr::fail(ctx, KHRR_EINVAL, "unsupported i_per_lane %d", t.i_per_lane);
  // Persistent grid: at most 8 waves per CU in flight, never more waves than items.
  const long long waves = std::min<long long>(n_items, 8ll * ctx->cu_count);
  const int wpb = t.waves_per_block;
  int* counter = ctx->dev_words;
  khrr::Timed tm(ctx, KHRR_K_GRAVITY);
  hipLaunchKernelGGL(set_word, dim3(1), dim3(1), 0, ctx->stream, counter, 0);
  fn(dim3(khrr::ceil_div(waves, wpb)), dim3(64 * wpb), ctx->stream, n_items, items, sys_off, uni,
     n_total, pos, h, m, t.j_tile, counter, acc, phi, (flags & KHRR_ACCUMULATE) ? 1 : 0);
  return khrr::check_launch(ctx, "batched_kernel");
}

int check_composition(khrr_ctx* ctx, int n_sys, const int* host_off) {
  if (n_sys < 0) return khrr::fail(ctx, KHRR_EINVAL, "n_sys < 0");
  if (n_sys > 0 && !host_off) return khrr::fail(ctx, KHRR_EINVAL, "host_sys_off is NULL");
  if (n_sys > 0 && host_off[0] != 0) return khrr::fail(ctx, KHRR_EINVAL, "sys_off[0] must be 0");
  for (int s = 0; s < n_sys; ++s)
    if (host_off[s + 1] < host_off[s])
      return khrr::fail(ctx, KHRR_EINVAL, "sys_off must be non-decreasing (at %d)", s);
  return 0;
}

// Deterministic synthetic bodies for tuning: a hashed cube.
__global__ void synth_bodies(int n, float* __restrict__ pos, float* __restrict__ m,
                             float* __restrict__ h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < 3; ++c) {
    unsigned x = (unsigned)i * 2654435761u + 0x9E3779B9u * (c + 1);
    x ^= x >> 15;
    x *= 0x2C1B3C6Du;
    x ^= x >> 12;
    pos[(size_t)c * n + i] = (float)(x & 0xFFFFFF) / 16777216.0f * 2.f - 1.f;
  }
  m[i] = 1.f / n;
  h[i] = 8e-6f;
}

template <typename F>
int time_candidate(khrr_ctx* ctx, F&& run, double* ms) {
  hipEvent_t a, b;
  KHRR_HIP(ctx, hipEventCreate(&a));
  KHRR_HIP(ctx, hipEventCreate(&b));
  int rc = run();  // warm-up
  if (!rc) {
    (void)hipEventRecord(a, ctx->stream);
    rc = run();
    (void)hipEventRecord(b, ctx->stream);
  }
  float t = 0.f;
  if (!rc && hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&t, a, b) == hipSuccess)
    *ms = t;
  else if (!rc)
    rc = khrr::fail(ctx, KHRR_EDEVICE, "timing failed");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

}  // namespace

extern "C" {

int khrr_accel_batched(khrr_ctx* ctx, int n_sys, const int* sys_off, const int* host_sys_off,
                       const float* pos, const float* h, const float* m, float* acc, float* phi,
                       unsigned flags, const khrr_tune* tune) {
  KHRR_TRY(khrr::enter(ctx));
  KHRR_TRY(check_composition(ctx, n_sys, host_sys_off));
  if (flags & ~(KHRR_ACCUMULATE | KHRR_COMPENSATED))
    return khrr::fail(ctx, KHRR_EINVAL, "unknown flags 0x%x", flags);
  if (!acc && !phi) return khrr::fail(ctx, KHRR_EINVAL, "acc and phi are both NULL");
  if (n_sys > 0 && host_sys_off[n_sys] > 0 && (!sys_off || !pos || !h || !m))
    return khrr::fail(ctx, KHRR_EINVAL, "null input pointer");
  khrr_tune t;
  if (tune)
    t = *tune;
  else if (ctx->has_batch_tuned)
    t = ctx->batch_tuned;
  else
    t = khrr_tune{4, 4, 1, 256};
  t.j_split = 1;
  KHRR_TRY(khrr::validate_tune(ctx, t));
  return batched_launch(ctx, n_sys, sys_off, host_sys_off, pos, h, m, acc, phi, flags, t);
}

int khrr_tune_gravity(khrr_ctx* ctx, int n, khrr_tune* best, double* best_ms) {
  KHRR_TRY(khrr::enter(ctx));
  if (n < 1 || !best) return khrr::fail(ctx, KHRR_EINVAL, "bad arguments");
  // Timing runs use the full i-set and at most 65536 j (the launch shape
  // depends on the i-groups and the wave count, not on the j length).
  const int nj = std::min(n, 65536);
  float *pos = nullptr, *m = nullptr, *h = nullptr, *acc = nullptr;
  KHRR_HIP(ctx, hipMalloc(&pos, 3ull * n * sizeof(float)));
  int rc = 0;
  if (hipMalloc(&m, (size_t)n * sizeof(float)) != hipSuccess ||
      hipMalloc(&h, (size_t)n * sizeof(float)) != hipSuccess ||
      hipMalloc(&acc, 3ull * n * sizeof(float)) != hipSuccess)
    rc = khrr::fail(ctx, KHRR_ENOMEM, "tuning buffers");
  if (!rc) {
    hipLaunchKernelGGL(synth_bodies, dim3(khrr::ceil_div(n, 256)), dim3(256), 0, ctx->stream, n,
                       pos, m, h);
    rc = khrr::check_launch(ctx, "synth_bodies");
  }
  double best_t = 1e300;
  khrr_tune best_c = khrr::default_tune(ctx, n, n);
  const int ipls[] = {1, 2, 4, 8};
  const int wpbs[] = {1, 2, 4};
  const int splits[] = {1, 2, 4, 8, 16};
  const bool timing = ctx->timing;
  ctx->timing = false;
  for (int ipl : ipls) {
    for (int wpb : wpbs) {
      for (int sp : splits) {
        if (rc) break;
        if (sp > 1 && (long long)nj / sp < 1024) continue;
        khrr_tune c{ipl, wpb, sp, 256, 0};
        double ms = 0.0;
        rc = time_candidate(ctx, [&] {
          return khrr::accel_launch(ctx, n, pos, h, nj, pos, m, h, acc, nullptr, 0, c);
        }, &ms);
        if (!rc && ms < best_t) {
          best_t = ms;
          best_c = c;
        }
      }
    }
  }
  // Pair-symmetric candidates run on the full system (i-set == j-set) and
  // are compared per pair-interaction with the one-sided ones.
  double best_per_pair = best_t / ((double)n * nj);
  for (int ipl : {4, 8}) {
    if (rc || !khrr::sym_eligible(n, ipl)) break;
    khrr_tune c{ipl, best_c.waves_per_block, best_c.j_split, 256, 1};
    double ms = 0.0;
    rc = time_candidate(ctx, [&] {
      return khrr::accel_launch(ctx, n, pos, h, n, pos, m, h, acc, nullptr, 0, c);
    }, &ms);
    if (!rc && ms / ((double)n * n) < best_per_pair) {
      best_per_pair = ms / ((double)n * n);
      best_t = ms;
      best_c = c;
    }
  }
  ctx->timing = timing;
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(pos);
  if (m) (void)hipFree(m);
  if (h) (void)hipFree(h);
  if (acc) (void)hipFree(acc);
  if (rc) return rc;
  khrr::store_tune(ctx, n, best_c);
  *best = best_c;
  if (best_ms) *best_ms = best_t;
  return 0;
}

int khrr_tune_batched(khrr_ctx* ctx, int n_sys, const int* host_sys_off, khrr_tune* best,
                      double* best_ms) {
  KHRR_TRY(khrr::enter(ctx));
  if (!best) return khrr::fail(ctx, KHRR_EINVAL, "best is NULL");
  KHRR_TRY(check_composition(ctx, n_sys, host_sys_off));
  const int n = n_sys > 0 ? host_sys_off[n_sys] : 0;
  if (n == 0) return khrr::fail(ctx, KHRR_EINVAL, "empty composition");
  float *pos = nullptr, *m = nullptr, *h = nullptr, *acc = nullptr;
  int* off = nullptr;
  int rc = 0;
  if (hipMalloc(&pos, 3ull * n * sizeof(float)) != hipSuccess ||
      hipMalloc(&m, (size_t)n * sizeof(float)) != hipSuccess ||
      hipMalloc(&h, (size_t)n * sizeof(float)) != hipSuccess ||
      hipMalloc(&acc, 3ull * n * sizeof(float)) != hipSuccess ||
      hipMalloc(&off, (n_sys + 1) * sizeof(int)) != hipSuccess)
    rc = khrr::fail(ctx, KHRR_ENOMEM, "tuning buffers");
  if (!rc) {
    hipLaunchKernelGGL(synth_bodies, dim3(khrr::ceil_div(n, 256)), dim3(256), 0, ctx->stream, n,
                       pos, m, h);
    rc = khrr::check_launch(ctx, "synth_bodies");
  }
  if (!rc && hipMemcpy(off, host_sys_off, (n_sys + 1) * sizeof(int), hipMemcpyHostToDevice) !=
                 hipSuccess)
    rc = khrr::fail(ctx, KHRR_EDEVICE, "offset upload failed");
  double best_t = 1e300;
  khrr_tune best_c{4, 4, 1, 256};
  const bool timing = ctx->timing;
  ctx->timing = false;
  for (int ipl : {2, 4, 8}) {
    for (int wpb : {1, 2, 4}) {
      if (rc) break;
      khrr_tune c{ipl, wpb, 1, 256};
      double ms = 0.0;
      rc = time_candidate(ctx, [&] {
        return batched_launch(ctx, n_sys, off, host_sys_off, pos, h, m, acc, nullptr, 0, c);
      }, &ms);
      if (!rc && ms < best_t) {
        best_t = ms;
        best_c = c;
      }
    }
  }
  ctx->timing = timing;
  (void)hipStreamSynchronize(ctx->stream);
  for (void* p : {(void*)pos, (void*)m, (void*)h, (void*)acc, (void*)off})
    if (p) (void)hipFree(p);
  // The cached work list points at the tuning composition's i-groups only
  // through batch_off, which stays valid; nothing else to reset.
  if (rc) return rc;
  ctx->batch_tuned = best_c;
  ctx->has_batch_tuned = true;
  *best = best_c;
  if (best_ms) *best_ms = best_t;
  return 0;
}

}  // extern "C"
