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
#include <type_traits>

#include "khrr_internal.h"
#include "khrr_pair.h"

namespace {

using khrr::Accum;

// One work item: i-group `group` of a system whose bodies are
// [begin, begin + count) of planes with stride `ld`, against the system's
// j-range [j0, j1); its first j-tile's flag is tile_flags[flag_base].  A
// system whose j-range is cut into k > 1 pieces writes partial sums for
// piece `split` at part[NC*unit + (split*NC + c)*count + local i]
// (split < 0: the item covers the whole j-range and writes the outputs).
struct Item {
  int group, begin, count, flag_base, j0, j1, unit, split;
};
// One cut system for the fixed-order reduction of its k partial sums.
struct RedItem {
  int begin, count, k, unit;
};

// Number of j-pieces of a system of n bodies: a rule of the system's own
// size, so the summation order never depends on the rest of the batch.
// Pieces of about 1024 j-bodies keep a wave's item short (2048 serial pair
// steps per lane at i_per_lane 2), so a round of a few small systems is not
// held up by a few long waves: such a round lasts about one item.
inline int pieces_of(int n) { return std::max(1, std::min(32, (n + 1023) / 1024)); }

template <int IPL, bool DO_ACC, bool DO_PHI, bool COMP>
__global__ __launch_bounds__(256) void batched_kernel(
    int n_items, const Item* __restrict__ items, const int* __restrict__ tile_flags, int ld,
    const float* __restrict__ pos, const float* __restrict__ hh_all, const float* __restrict__ m,
    int j_tile, int* __restrict__ counter, float* __restrict__ acc, float* __restrict__ phi,
    float* __restrict__ part, int accumulate) {
  constexpr int NC = COMP ? 8 : 4;
  const int lane = threadIdx.x & (khrr::kWave - 1);
  const float* xa = pos;
  const float* ya = pos + ld;
  const float* za = pos + 2 * (size_t)ld;
  for (;;) {
    int got = 0;
    if (lane == 0) got = atomicAdd(counter, 1);
    const int item = __builtin_amdgcn_readfirstlane(got);
    if (item >= n_items) break;
    const Item it = items[item];
    const int base = it.begin;
    const int n = it.count;

    float xi[IPL], yi[IPL], zi[IPL], hi[IPL];
    int loc[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
      loc[k] = it.group * (IPL * khrr::kWave) + k * khrr::kWave + lane;
      const bool ok = loc[k] < n;
      const int i = base + loc[k];
      xi[k] = ok ? xa[i] : 0.f;
      yi[k] = ok ? ya[i] : 0.f;
      zi[k] = ok ? za[i] : 0.f;
      hi[k] = ok ? hh_all[i] : 1.f;
    }
    Accum<IPL> a;
    a.zero();
    for (int jt = it.j0, t = it.flag_base + (it.j0 - base) / j_tile; jt < it.j1;
         jt += j_tile, ++t) {
      const int je = min(it.j1, jt + j_tile);
      khrr::tile_step<IPL, DO_ACC, DO_PHI, COMP>(jt, je, tile_flags[t], xa, ya, za, m, hh_all, xi,
                                                 yi, zi, hi, a);
    }
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
      if (loc[k] >= n) continue;
      if (it.split < 0) {
        khrr::store_final<IPL, DO_ACC, DO_PHI, COMP>(a, k, base + loc[k], (size_t)ld, acc, phi,
                                                     accumulate);
        continue;
      }
      float* p = part + (size_t)NC * it.unit + (size_t)it.split * NC * n + loc[k];
      const size_t N = (size_t)n;
      if (DO_ACC) {
        p[0] = a.ax[k];
        p[N] = a.ay[k];
        p[2 * N] = a.az[k];
      }
      if (DO_PHI) p[3 * N] = a.ap[k];
      if (COMP) {
        if (DO_ACC) {
          p[4 * N] = a.lx[k];
          p[5 * N] = a.ly[k];
          p[6 * N] = a.lz[k];
        }
        if (DO_PHI) p[7 * N] = a.lp[k];
      }
    }
  }
}

// Adds the k partial sums of each cut system in piece order (the order of
// the plain kernel's split reduction) and writes the outputs.
template <bool COMP>
__global__ __launch_bounds__(256) void batched_reduce(const RedItem* __restrict__ red, int ld,
                                                      int do_acc, int do_phi,
                                                      const float* __restrict__ part,
                                                      float* __restrict__ acc,
                                                      float* __restrict__ phi, int accumulate) {
  constexpr int NC = COMP ? 8 : 4;
  const RedItem r = red[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= r.count) return;
  const size_t N = (size_t)r.count;
  const float* p0 = part + (size_t)NC * r.unit;
  for (int c = 0; c < 4; ++c) {
    if ((c < 3 && !do_acc) || (c == 3 && !do_phi)) continue;
    float hi = p0[c * N + i];
    float lo = COMP ? p0[(c + 4) * N + i] : 0.f;
    for (int s = 1; s < r.k; ++s) {
      const float* p = p0 + (size_t)s * NC * N;
      khrr::fold(hi, lo, p[c * N + i], COMP);
      if (COMP) lo += p[(c + 4) * N + i];
    }
    const float v = COMP ? hi + lo : hi;
    const size_t g = (size_t)r.begin + i;
    if (c < 3) {
      float* a = acc + (size_t)c * ld + g;
      *a = accumulate ? *a + v : v;
    } else {
      phi[g] = accumulate ? phi[g] - v : -v;
    }
  }
}

// tile_flags[t]: kUniformH / kUniformM when all h / all m of j-tile t (a
// range [tiles[t].x, tiles[t].y) inside one system) are bitwise equal.  The
// tiles of a system start at its first body, exactly as the plain kernel's
// tiles start at j = 0, so both forms take the same arithmetic path.
// It also clears the persistent grid's item counter, which every launch
// sequence runs after it.
__global__ __launch_bounds__(256) void tile_flags_kernel(int n_tiles, const int2* __restrict__ tiles,
                                                         const float* __restrict__ h,
                                                         const float* __restrict__ m,
                                                         int* __restrict__ flags,
                                                         int* __restrict__ counter) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0;
  const int lane = threadIdx.x & (khrr::kWave - 1);
  const int t = blockIdx.x * (blockDim.x / khrr::kWave) + threadIdx.x / khrr::kWave;
  if (t >= n_tiles) return;
  const int b = tiles[t].x, e = tiles[t].y;
  const unsigned href = __float_as_uint(h[b]);
  const unsigned mref = __float_as_uint(m[b]);
  bool same_h = true, same_m = true;
  for (int j = b + lane; j < e; j += khrr::kWave) {
    same_h &= __float_as_uint(h[j]) == href;
    same_m &= __float_as_uint(m[j]) == mref;
  }
  const unsigned long long bad_h = __ballot(!same_h);
  const unsigned long long bad_m = __ballot(!same_m);
  if (lane == 0)
    flags[t] = (bad_h == 0ull ? khrr::kUniformH : 0) | (bad_m == 0ull ? khrr::kUniformM : 0);
}

using BatchLauncher = void (*)(dim3, dim3, hipStream_t, int, const Item*, const int*, int,
                               const float*, const float*, const float*, int, int*, float*, float*,
                               float*, int);

template <int IPL, bool A, bool P, bool C>
void launch_batched(dim3 g, dim3 b, hipStream_t st, int n_items, const Item* items,
                    const int* flags, int ld, const float* pos, const float* h, const float* m,
                    int j_tile, int* counter, float* acc, float* phi, float* part, int accum) {
  hipLaunchKernelGGL((batched_kernel<IPL, A, P, C>), g, b, 0, st, n_items, items, flags, ld, pos, h,
                     m, j_tile, counter, acc, phi, part, accum);
}

template <int IPL>
BatchLauncher pick_batched_mode(bool a, bool p, bool c) {
  if (a && p) return c ? launch_batched<IPL, true, true, true> : launch_batched<IPL, true, true, false>;
  if (a) return c ? launch_batched<IPL, true, false, true> : launch_batched<IPL, true, false, false>;
  return c ? launch_batched<IPL, false, true, true> : launch_batched<IPL, false, true, false>;
}

template <int IPL>
const void* pick_kernel_mode(bool a, bool p, bool c) {
  if (a && p)
    return c ? (const void*)batched_kernel<IPL, true, true, true>
             : (const void*)batched_kernel<IPL, true, true, false>;
  if (a)
    return c ? (const void*)batched_kernel<IPL, true, false, true>
             : (const void*)batched_kernel<IPL, true, false, false>;
  return c ? (const void*)batched_kernel<IPL, false, true, true>
           : (const void*)batched_kernel<IPL, false, true, false>;
}

const void* pick_kernel(int ipl, bool a, bool p, bool c) {
  switch (ipl) {
    case 1: return pick_kernel_mode<1>(a, p, c);
    case 2: return pick_kernel_mode<2>(a, p, c);
    case 4: return pick_kernel_mode<4>(a, p, c);
    case 8: return pick_kernel_mode<8>(a, p, c);
    default: return nullptr;
  }
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

// Systems of at least 16384 bodies take the pair-symmetric path (16 or
// more superblocks of 1024, 2048 or 4096 bodies: the one-sided diagonal
// items are at most 1/16 of the system's work).  Like the piece count, a
// rule of the system's size.
inline bool on_sym_path(int n) { return n >= 16384 && khrr::sym_eligible(n, 8, 0); }

// Device work list of one composition.
struct WorkList {
  Item* items = nullptr;
  int n_items = 0;
  int* tile_flags = nullptr;
  RedItem* red = nullptr;
  int n_red = 0;
  int max_red_count = 0;
  long long part_units = 0;  // sum of k*count over cut systems
  khrr::SymItem* sym_items = nullptr;
  int n_sym_items = 0;
  int n_sym_big = 0;
  int n_sym_mid = 0;
  khrr::SymSys* sym_sys = nullptr;
  int n_sym_sys = 0;
  int max_sym_n = 0;
  long long sym_floats = 0;
  const int2* tiles = nullptr;
  int n_tiles = 0;
  size_t dev_bytes = 0;  // size of the device block the pointers live in
};

// Byte offsets of the tables in the device block (and the pinned copy).
struct Layout {
  size_t items, tiles, red, sym_items, sym_sys, flags, host_bytes, dev_bytes;
  Layout(size_t n_items, size_t n_tiles, size_t n_red, size_t n_sym_items, size_t n_sym_sys) {
    auto need = [](size_t c, size_t e) { return khrr::Carver::need(c > 0 ? c : 1, e); };
    items = 0;
    tiles = items + need(n_items, sizeof(Item));
    red = tiles + need(n_tiles, sizeof(int2));
    sym_items = red + need(n_red, sizeof(RedItem));
    sym_sys = sym_items + need(n_sym_items, sizeof(khrr::SymItem));
    flags = sym_sys + need(n_sym_sys, sizeof(khrr::SymSys));
    host_bytes = flags;
    dev_bytes = flags + need(n_tiles, sizeof(int));
  }
};

// Builds (or reuses) the device work list for this composition: one-sided
// items longest system first, the flag ranges of every system (j-tiles, or
// superblocks on the pair-symmetric path), the cut systems' reduction
// list, the pair-symmetric items and systems.  The flags themselves are
// recomputed by every launch (masses and softening may change).
int prepare_items(khrr_ctx* ctx, int n_sys, const int* host_begin, const int* host_count, int ipl,
                  int j_tile, WorkList* wl) {
  std::vector<int> key(2 * (size_t)n_sys + 1);
  for (int s = 0; s < n_sys; ++s) {
    key[2 * s] = host_begin[s];
    key[2 * s + 1] = host_count[s];
  }
  key[2 * (size_t)n_sys] = j_tile;
  const bool same = ctx->batch_ipl == ipl && ctx->batch_off == key;
  if (!same) {
    std::vector<int> flag_base(n_sys);
    std::vector<int2> tiles;
    std::vector<khrr::SymItem> sym_big, sym_mid, sym_small;
    std::vector<khrr::SymSys> sym_sys;
    long long sym_floats = 0;
    int max_sym = 0;
    for (int s = 0; s < n_sys; ++s) {
      flag_base[s] = (int)tiles.size();
      const int b = host_begin[s], n = host_count[s], e = b + n;
      const int SB = on_sym_path(n) ? khrr::sym_superblock(n, 0) : 0;
      const int step = SB ? SB : j_tile;
      for (int jt = b; jt < e; jt += step) tiles.push_back(make_int2(jt, std::min(e, jt + step)));
      if (!SB) continue;
      const int n_sb = (n + SB - 1) / SB;
      auto& dst = SB == 4096 ? sym_big : SB == 2048 ? sym_mid : sym_small;
      for (int A = 0; A < n_sb; ++A)
        for (int B = A; B < n_sb; ++B)
          dst.push_back(khrr::SymItem{b, n, A, B, flag_base[s], 0, sym_floats});
      sym_sys.push_back(khrr::SymSys{b, n, n_sb, 0, sym_floats});
      sym_floats += 4ll * n_sb * n;
      max_sym = std::max(max_sym, n);
    }
    std::vector<khrr::SymItem> sym_items(sym_big);
    sym_items.insert(sym_items.end(), sym_mid.begin(), sym_mid.end());
    sym_items.insert(sym_items.end(), sym_small.begin(), sym_small.end());
    std::vector<int> order(n_sys);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return host_count[a] > host_count[b]; });
    std::vector<Item> list;
    std::vector<RedItem> red;
    long long unit = 0;
    int max_red = 0;
    for (int s : order) {
      const int n = host_count[s], b = host_begin[s];
      if (n == 0 || on_sym_path(n)) continue;
      const int groups = (int)khrr::ceil_div(n, 64ll * ipl);
      // pieces as khrr_accel's j_split: chunk rounded up to whole tiles
      long long chunk = ((long long)n + pieces_of(n) - 1) / pieces_of(n);
      chunk = (chunk + j_tile - 1) / j_tile * j_tile;
      const int k = (int)khrr::ceil_div(n, chunk);
      for (int piece = 0; piece < k; ++piece) {
        const int j0 = b + (int)(piece * chunk);
        const int j1 = std::min(b + n, (int)(j0 + chunk));
        for (int g = 0; g < groups; ++g)
          list.push_back(Item{g, b, n, flag_base[s], j0, j1, k > 1 ? (int)unit : 0,
                              k > 1 ? piece : -1});
      }
      if (k > 1) {
        red.push_back(RedItem{b, n, k, (int)unit});
        unit += (long long)k * n;
        max_red = std::max(max_red, n);
      }
    }
    if (unit > INT32_MAX || list.size() > INT32_MAX || sym_items.size() > INT32_MAX ||
        tiles.size() > INT32_MAX)
      return khrr::fail(ctx, KHRR_EINVAL, "batch too large");
    const Layout L(list.size(), tiles.size(), red.size(), sym_items.size(), sym_sys.size());
    KHRR_HIP(ctx, hipStreamSynchronize(ctx->stream));  // old list/pinned copy may be in flight
    if (L.dev_bytes > ctx->batch_dev_bytes) {
      if (ctx->batch_dev) KHRR_HIP(ctx, hipFree(ctx->batch_dev));
      ctx->batch_dev = nullptr;
      ctx->batch_dev_bytes = 0;
      KHRR_HIP(ctx, hipMalloc(&ctx->batch_dev, L.dev_bytes));
      ctx->batch_dev_bytes = L.dev_bytes;
    }
    if (L.host_bytes > ctx->batch_pinned_bytes) {
      if (ctx->batch_pinned) KHRR_HIP(ctx, hipHostFree(ctx->batch_pinned));
      ctx->batch_pinned = nullptr;
      ctx->batch_pinned_bytes = 0;
      KHRR_HIP(ctx, hipHostMalloc(&ctx->batch_pinned, L.host_bytes));
      ctx->batch_pinned_bytes = L.host_bytes;
    }
    char* pin = static_cast<char*>(ctx->batch_pinned);
    auto put = [&](size_t off, const void* src, size_t bytes) {
      if (bytes) std::memcpy(pin + off, src, bytes);
    };
    put(L.items, list.data(), list.size() * sizeof(Item));
    put(L.tiles, tiles.data(), tiles.size() * sizeof(int2));
    put(L.red, red.data(), red.size() * sizeof(RedItem));
    put(L.sym_items, sym_items.data(), sym_items.size() * sizeof(khrr::SymItem));
    put(L.sym_sys, sym_sys.data(), sym_sys.size() * sizeof(khrr::SymSys));
    KHRR_HIP(ctx, hipMemcpyAsync(ctx->batch_dev, pin, L.host_bytes, hipMemcpyHostToDevice,
                                 ctx->stream));
    ctx->batch_off = key;
    ctx->batch_ipl = ipl;
    ctx->batch_items = (int)list.size();
    ctx->batch_tiles = (int)tiles.size();
    ctx->batch_red = (int)red.size();
    ctx->batch_red_max = max_red;
    ctx->batch_units = unit;
    ctx->batch_sym_items = (int)sym_items.size();
    ctx->batch_sym_big = (int)sym_big.size();
    ctx->batch_sym_mid = (int)sym_mid.size();
    ctx->batch_sym_sys = (int)sym_sys.size();
    ctx->batch_sym_max = max_sym;
    ctx->batch_sym_floats = sym_floats;
  }
  const Layout L(ctx->batch_items, ctx->batch_tiles, ctx->batch_red, ctx->batch_sym_items,
                 ctx->batch_sym_sys);
  char* dev = static_cast<char*>(ctx->batch_dev);
  wl->items = reinterpret_cast<Item*>(dev + L.items);
  const int2* tiles = reinterpret_cast<const int2*>(dev + L.tiles);
  wl->red = reinterpret_cast<RedItem*>(dev + L.red);
  wl->sym_items = reinterpret_cast<khrr::SymItem*>(dev + L.sym_items);
  wl->sym_sys = reinterpret_cast<khrr::SymSys*>(dev + L.sym_sys);
  wl->tile_flags = reinterpret_cast<int*>(dev + L.flags);
  wl->n_items = ctx->batch_items;
  wl->n_red = ctx->batch_red;
  wl->max_red_count = ctx->batch_red_max;
  wl->part_units = ctx->batch_units;
  wl->n_sym_items = ctx->batch_sym_items;
  wl->n_sym_big = ctx->batch_sym_big;
  wl->n_sym_mid = ctx->batch_sym_mid;
  wl->n_sym_sys = ctx->batch_sym_sys;
  wl->max_sym_n = ctx->batch_sym_max;
  wl->sym_floats = ctx->batch_sym_floats;
  wl->tiles = tiles;
  wl->n_tiles = ctx->batch_tiles;
  wl->dev_bytes = L.dev_bytes;
  return 0;
}

// The same work list with every pointer moved from block `from` to `to`.
WorkList rebase(const WorkList& w, const void* from, void* to) {
  auto mv = [&](auto* p) {
    using T = std::remove_reference_t<decltype(*p)>;
    return p ? reinterpret_cast<T*>(static_cast<char*>(to) +
                                    (reinterpret_cast<const char*>(p) -
                                     static_cast<const char*>(from)))
             : p;
  };
  WorkList r = w;
  r.items = mv(w.items);
  r.tile_flags = mv(w.tile_flags);
  r.red = mv(w.red);
  r.sym_items = mv(w.sym_items);
  r.sym_sys = mv(w.sym_sys);
  r.tiles = mv(w.tiles);
  return r;
}

// Resident blocks per CU of the persistent kernel for this shape (cached).
int blocks_per_cu(int ipl, bool a, bool p, bool c, int wpb) {
  static int cache[4][3][2][5] = {};
  const int li = ipl == 1 ? 0 : ipl == 2 ? 1 : ipl == 4 ? 2 : 3;
  const int mi = a && p ? 0 : a ? 1 : 2;
  int& v = cache[li][mi][c][wpb];
  if (v == 0) {
    int nb = 0;
    const void* f = pick_kernel(ipl, a, p, c);
    if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * wpb, 0) != hipSuccess ||
        nb < 1)
      nb = std::max(1, 8 / wpb);
    v = nb;
  }
  return v;
}

// The persistent one-sided grid over the work list's (i-group, j-piece)
// items on stream `st`: as many blocks as stay resident, never more waves
// than items; then the piece reduction of the cut systems.
int launch_one_sided(khrr_ctx* ctx, hipStream_t st, const WorkList& wl, int ld, const float* pos,
                     const float* h, const float* m, float* acc, float* phi, bool comp,
                     int accumulate, const khrr_tune& t, float* part) {
  using namespace khrr;
  BatchLauncher fn = pick_batched(t.i_per_lane, acc != nullptr, phi != nullptr, comp);
  const int wpb = t.waves_per_block;
  const long long resident =
      (long long)blocks_per_cu(t.i_per_lane, acc != nullptr, phi != nullptr, comp, wpb) *
      ctx->cu_count;
  const long long blocks = std::min<long long>(resident, ceil_div(wl.n_items, wpb));
  int* counter = ctx->dev_words;  // cleared by tile_flags_kernel
  fn(dim3((unsigned)blocks), dim3(64 * wpb), st, wl.n_items, wl.items, wl.tile_flags, ld, pos, h,
     m, t.j_tile, counter, acc, phi, part, accumulate);
  if (wl.n_red > 0) {
    const dim3 grid(ceil_div(wl.max_red_count, 256), wl.n_red);
    if (comp)
      hipLaunchKernelGGL(batched_reduce<true>, grid, dim3(256), 0, st, wl.red, ld,
                         (int)(acc != nullptr), (int)(phi != nullptr), part, acc, phi, accumulate);
    else
      hipLaunchKernelGGL(batched_reduce<false>, grid, dim3(256), 0, st, wl.red, ld,
                         (int)(acc != nullptr), (int)(phi != nullptr), part, acc, phi, accumulate);
  }
  return check_launch(ctx, "batched_kernel");
}

// The launches of one batched evaluation over a prepared work list: tile
// flags, then the pair-symmetric items with their slot reduction and the
// one-sided grid with its piece reduction.  The two parts touch disjoint
// systems, so when both are present the one-sided part runs on the
// context's second stream (forked after the flags, joined at the end) and
// fills the CUs the symmetric launch leaves idle near its end.  No host
// synchronisation and no allocation, so the sequence can be captured into
// a graph (the fork and join become graph edges).
int launch_work(khrr_ctx* ctx, const WorkList& wl, int ld, const float* pos, const float* h,
                const float* m, float* acc, float* phi, unsigned flags, const khrr_tune& t,
                float* part, float* sym_part) {
  using namespace khrr;
  const bool comp = (flags & KHRR_COMPENSATED) != 0;
  const int accumulate = (flags & KHRR_ACCUMULATE) ? 1 : 0;
  if (wl.n_tiles > 0)  // every system with items has tiles
    hipLaunchKernelGGL(tile_flags_kernel, dim3(ceil_div(wl.n_tiles, 4)), dim3(256), 0, ctx->stream,
                       wl.n_tiles, wl.tiles, h, m, wl.tile_flags, ctx->dev_words);
  KHRR_TRY(check_launch(ctx, "tile_flags_kernel"));
  const bool fork = wl.n_items > 0 && wl.n_sym_items > 0;
  if (fork) {
    KHRR_TRY(aux_stream(ctx));
    KHRR_HIP(ctx, hipEventRecord(ctx->fork_event, ctx->stream));
    KHRR_HIP(ctx, hipStreamWaitEvent(ctx->aux_stream, ctx->fork_event, 0));
    KHRR_TRY(launch_one_sided(ctx, ctx->aux_stream, wl, ld, pos, h, m, acc, phi, comp, accumulate,
                              t, part));
    KHRR_HIP(ctx, hipEventRecord(ctx->join_event, ctx->aux_stream));
  }
  KHRR_TRY(sym_items_launch(ctx, wl.n_sym_big, wl.n_sym_mid,
                            wl.n_sym_items - wl.n_sym_big - wl.n_sym_mid, wl.sym_items,
                            wl.n_sym_sys, wl.sym_sys, wl.max_sym_n, ld, pos, h, m, wl.tile_flags,
                            sym_part, acc, phi, flags));
  if (fork) {
    KHRR_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->join_event, 0));
    return 0;
  }
  if (wl.n_items == 0) return 0;
  return launch_one_sided(ctx, ctx->stream, wl, ld, pos, h, m, acc, phi, comp, accumulate, t,
                          part);
}

}  // namespace

namespace khrr {

int round_ipl(const khrr_ctx* ctx, int n_sys, const int* count, int pref) {
  const long long want = 2ll * 4 * ctx->cu_count;  // two items per SIMD
  bool any = false;
  for (int ipl = pref; ipl >= 1; ipl /= 2) {
    long long items = 0;
    for (int s = 0; s < n_sys; ++s) {
      const int n = count[s];
      if (n <= 0 || on_sym_path(n)) continue;
      any = true;
      items += (long long)ceil_div(n, 64ll * ipl) * pieces_of(n);
    }
    if (!any || items >= want) return ipl;
  }
  return 1;
}

size_t batched_part_bytes(int n, bool compensated) {
  if (n <= 0) return 0;
  if (on_sym_path(n)) {
    const int sb = sym_superblock(n, 0);
    const size_t n_sb = (size_t)(n + sb - 1) / sb;
    return n_sb * 4 * (size_t)n * sizeof(float);
  }
  const int k = pieces_of(n);  // the piece count before tile rounding bounds it
  return k > 1 ? (size_t)k * (compensated ? 8 : 4) * (size_t)n * sizeof(float) : 0;
}

int batched_segments(khrr_ctx* ctx, int n_sys, const int* host_begin, const int* host_count,
                     int ld, const float* pos, const float* h, const float* m, float* acc,
                     float* phi, unsigned flags, const khrr_tune& t) {
  long long total = 0;
  for (int s = 0; s < n_sys; ++s) total += host_count[s];
  if (total == 0 || n_sys == 0) return 0;
  WorkList wl;
  KHRR_TRY(prepare_items(ctx, n_sys, host_begin, host_count, t.i_per_lane, t.j_tile, &wl));
  const bool comp = (flags & KHRR_COMPENSATED) != 0;
  if (!pick_batched(t.i_per_lane, acc != nullptr, phi != nullptr, comp))
    return fail(ctx, KHRR_EINVAL, "unsupported i_per_lane %d", t.i_per_lane);
  float* part = nullptr;
  float* sym_part = nullptr;
  if (wl.n_red > 0 || wl.n_sym_items > 0) {
    const size_t pb = Carver::need((size_t)(comp ? 8 : 4) * wl.part_units, sizeof(float));
    const size_t sb = Carver::need((size_t)wl.sym_floats, sizeof(float));
    void* base = nullptr;
    KHRR_TRY(scratch(ctx, pb + sb, &base));
    Carver cv(base);
    part = cv.take<float>((size_t)(comp ? 8 : 4) * wl.part_units);
    sym_part = cv.take<float>((size_t)wl.sym_floats);
  }
  const unsigned kf = flags & ~KHRR_GRAPH;
  if (!(flags & KHRR_GRAPH)) {
    Timed tm(ctx, KHRR_K_GRAVITY);
    return launch_work(ctx, wl, ld, pos, h, m, acc, phi, kf, t, part, sym_part);
  }

  // Graph path: the key is everything the captured launches depend on.
  std::vector<long long> key(ctx->batch_off.begin(), ctx->batch_off.end());
  for (long long v : {(long long)t.i_per_lane, (long long)t.waves_per_block, (long long)ld,
                      (long long)kf, (long long)(intptr_t)pos, (long long)(intptr_t)h,
                      (long long)(intptr_t)m, (long long)(intptr_t)acc, (long long)(intptr_t)phi,
                      (long long)(intptr_t)ctx->scratch, (long long)ctx->scratch_bytes})
    key.push_back(v);
  for (auto& g : ctx->graphs) {
    if (g.key != key) continue;
    g.last_use = ++ctx->graph_clock;
    Timed tm(ctx, KHRR_K_GRAVITY);
    if (g.failed) return launch_work(ctx, wl, ld, pos, h, m, acc, phi, kf, t, part, sym_part);
    KHRR_HIP(ctx, hipGraphLaunch(g.exec, ctx->stream));
    return 0;
  }
  // First call of this key: run it eagerly, then capture a replayable copy.
  {
    Timed tm(ctx, KHRR_K_GRAVITY);
    KHRR_TRY(launch_work(ctx, wl, ld, pos, h, m, acc, phi, kf, t, part, sym_part));
  }
  if (ctx->graphs.size() >= 4) {
    auto lru = std::min_element(ctx->graphs.begin(), ctx->graphs.end(),
                                [](const GraphEntry& a, const GraphEntry& b) {
                                  return a.last_use < b.last_use;
                                });
    KHRR_HIP(ctx, hipStreamSynchronize(ctx->stream));  // a replay may still run
    if (lru->exec) (void)hipGraphExecDestroy(lru->exec);
    if (lru->graph) (void)hipGraphDestroy(lru->graph);
    if (lru->dev) (void)hipFree(lru->dev);
    ctx->graphs.erase(lru);
  }
  GraphEntry g;
  g.key = key;
  g.last_use = ++ctx->graph_clock;
  bool ok = hipMalloc(&g.dev, wl.dev_bytes) == hipSuccess;
  if (ok && !ctx->capture_stream)
    ok = hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking) == hipSuccess;
  if (ok)  // the work list as it is now, ordered after the eager run's use of it
    ok = hipMemcpyAsync(g.dev, ctx->batch_dev, wl.dev_bytes, hipMemcpyDeviceToDevice,
                        ctx->stream) == hipSuccess &&
         hipStreamSynchronize(ctx->stream) == hipSuccess;
  if (ok) {
    const WorkList wg = rebase(wl, ctx->batch_dev, g.dev);
    hipStream_t saved = ctx->stream;
    const bool timing = ctx->timing;
    ctx->stream = ctx->capture_stream;
    ctx->timing = false;
    ok = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    const int rc = ok ? launch_work(ctx, wg, ld, pos, h, m, acc, phi, kf, t, part, sym_part) : 0;
    hipGraph_t graph = nullptr;
    if (ok) ok = hipStreamEndCapture(ctx->stream, &graph) == hipSuccess && rc == 0 && graph;
    ctx->stream = saved;
    ctx->timing = timing;
    g.graph = graph;
    if (ok) ok = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess;
  }
  if (!ok) {  // keep the eager path for this key
    (void)hipGetLastError();
    ctx->err.clear();
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    if (g.dev) (void)hipFree(g.dev);
    g.exec = nullptr;
    g.graph = nullptr;
    g.dev = nullptr;
    g.failed = true;
  }
  ctx->graphs.push_back(std::move(g));
  return 0;
}

}  // namespace khrr

namespace {

int batched_launch(khrr_ctx* ctx, int n_sys, const int* host_sys_off, const float* pos,
                   const float* h, const float* m, float* acc, float* phi, unsigned flags,
                   const khrr_tune& t) {
  std::vector<int> begin(n_sys), count(n_sys);
  for (int s = 0; s < n_sys; ++s) {
    begin[s] = host_sys_off[s];
    count[s] = host_sys_off[s + 1] - host_sys_off[s];
  }
  return khrr::batched_segments(ctx, n_sys, begin.data(), count.data(),
                                n_sys > 0 ? host_sys_off[n_sys] : 0, pos, h, m, acc, phi, flags,
                                t);
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
  if (flags & ~(KHRR_ACCUMULATE | KHRR_COMPENSATED | KHRR_GRAPH))
    return khrr::fail(ctx, KHRR_EINVAL, "unknown flags 0x%x", flags);
  if (!acc && !phi) return khrr::fail(ctx, KHRR_EINVAL, "acc and phi are both NULL");
  if (n_sys > 0 && host_sys_off[n_sys] > 0 && (!pos || !h || !m))
    return khrr::fail(ctx, KHRR_EINVAL, "null input pointer");
  khrr_tune t;
  if (tune) {
    t = *tune;
  } else if (ctx->has_batch_tuned) {
    t = ctx->batch_tuned;
  } else {
    t = khrr_tune{4, 4, 1, 256, 0};
    std::vector<int> count(n_sys);
    for (int s = 0; s < n_sys; ++s) count[s] = host_sys_off[s + 1] - host_sys_off[s];
    t.i_per_lane = khrr::round_ipl(ctx, n_sys, count.data(), 4);
  }
  t.j_split = 1;
  t.superblock = 0;  // pair-symmetric systems follow the rule of their own n
  KHRR_TRY(khrr::validate_tune(ctx, t));
  (void)sys_off;  // the host offsets define the segments
  return batched_launch(ctx, n_sys, host_sys_off, pos, h, m, acc, phi, flags, t);
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
  const bool timing = ctx->timing;
  ctx->timing = false;
  auto try_one_sided = [&](const khrr_tune& c) {
    double ms = 0.0;
    rc = time_candidate(ctx, [&] {
      return khrr::accel_launch(ctx, n, pos, h, nj, pos, m, h, acc, nullptr, 0, c);
    }, &ms);
    if (!rc && ms < best_t) {
      best_t = ms;
      best_c = c;
    }
  };
  // One-sided: i-bodies per lane x waves per block x j-ranges with 256-body
  // tiles, then the j-tile of the fastest of them.
  for (int ipl : {1, 2, 4, 8})
    for (int wpb : {1, 2, 4})
      for (int sp : {1, 2, 4, 8, 16}) {
        if (rc) break;
        if (sp > 1 && (long long)nj / sp < 1024) continue;
        try_one_sided(khrr_tune{ipl, wpb, sp, 256, 0, 0});
      }
  const khrr_tune base = best_c;
  for (int jt : {128, 512, 1024}) {
    if (rc) break;
    if (jt > nj) continue;
    khrr_tune c = base;
    c.j_tile = jt;
    try_one_sided(c);
  }
  // Pair-symmetric candidates (i-bodies per lane x superblock size) run on
  // the full system (i-set == j-set) and are compared per pair-interaction
  // with the one-sided ones.
  double best_per_pair = best_t / ((double)n * nj);
  const int sym_shapes[][2] = {{4, 1024}, {4, 2048}, {8, 2048}, {4, 4096}, {8, 4096}};
  for (const auto& sh : sym_shapes) {
    if (rc) break;
    if (!khrr::sym_eligible(n, sh[0], sh[1])) continue;
    khrr_tune c{sh[0], best_c.waves_per_block, best_c.j_split, best_c.j_tile, 1, sh[1]};
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
  ctx->tuned_ms[khrr::tune_bucket(n)] = best_t;
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
  khrr_tune best_c{4, 4, 1, 256, 0, 0};
  const bool timing = ctx->timing;
  ctx->timing = false;
  auto try_shape = [&](const khrr_tune& c) {
    double ms = 0.0;
    rc = time_candidate(ctx, [&] {
      return batched_launch(ctx, n_sys, host_sys_off, pos, h, m, acc, nullptr, 0, c);
    }, &ms);
    if (!rc && ms < best_t) {
      best_t = ms;
      best_c = c;
    }
  };
  // i-bodies per lane x waves per block with 256-body j-tiles, then the
  // j-tile of the fastest (it sets the one-sided items' partial-sum tiles
  // and so their results; the pair-symmetric systems follow their rule)
  for (int ipl : {2, 4, 8})
    for (int wpb : {1, 2, 4}) {
      if (rc) break;
      try_shape(khrr_tune{ipl, wpb, 1, 256, 0, 0});
    }
  const khrr_tune base = best_c;
  for (int jt : {128, 512, 1024}) {
    if (rc) break;
    khrr_tune c = base;
    c.j_tile = jt;
    try_shape(c);
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
