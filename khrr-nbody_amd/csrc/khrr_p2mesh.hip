// khrr_p2mesh.hip -- deterministic particle-to-mesh binning (gfx950).
//
// Replaces the sequence convert_particles_to_grid (nbody/src/nbody.cu:139-181):
//   assignParticlesToCells     p2mesh-kernel.cuh:11-57
//   computePrefixSums          p2mesh-kernel.cuh:62-105
//   adjustGlobalPrefixSums     p2mesh-kernel.cuh:110-123
//   computeCellMassesUnsorted  p2mesh-kernel.cuh:126-137
//
// Pipeline (all integer/FP64, bit-exact with a sequential evaluation):
//  1. assign:  cell index per particle from FP64 positions, same arithmetic
//              as the reference (subtract, IEEE divide, truncate with GPU
//              cvt saturation, clamp; z clamped with size[1]; y stride sz).
//  2. sort:    stable LSD radix sort of (cell, particle index) by the
//              cell's 512-cell block (cell >> 9: 15 bits at 256^3, two
//              8-bit passes); per pass a tile histogram, one scan block per
//              digit row, and a scatter whose in-block ranks come from
//              64-lane ballot matching, so a block keeps index order.  The
//              sorted (cell, index) pairs are stored interleaved, one 8-byte
//              store per body.
//  3. cells:   one 512-thread block per 512-cell block of the reference.
//              Binary searches on the sorted keys give the number of
//              particles before the block (= the reference's block-granular
//              start_id) and through it (= maxPBC); an LDS histogram and
//              scan give per-cell counts and exact exclusive offsets; a
//              stable in-block rank by cell (ballot groups per wave, byte
//              counters per wave in LDS) gives id_in_cell, the index-order
//              rank, and each cell adds its particles' FP64 masses in index
//              order from LDS.
// Every cell is written exactly once, so no output needs clearing first.
#include <algorithm>
#include <climits>

#include "khrr_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortThreads = 256;  // == kRadix: one digit per thread
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kCellBlock = KHRR_BIN_BLOCK;  // 512 cells, BLOCK_SIZE of common.cuh:8
constexpr int kBlockShift = 9;               // cell >> 9 is the cell's 512-cell block
static_assert((1 << kBlockShift) == kCellBlock, "block shift");

struct GridArgs {
  double min0, min1, min2;
  double cell0, cell1, cell2;
  int s0, s1, s2;
};

// (int) conversion as GPU cvt does it: truncate toward zero, saturate to
// the int range, NaN -> 0.
__device__ __forceinline__ int sat_trunc(double v) {
  if (v != v) return 0;
  if (v >= 2147483648.0) return INT_MAX;
  if (v <= -2147483649.0) return INT_MIN;
  return (int)v;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// The reference's cell index of one position (p2mesh-kernel.cuh:25-43).
__device__ __forceinline__ int cell_of(double x, double y, double z, const GridArgs& g) {
  const int ix = clampi(sat_trunc((x - g.min0) / g.cell0), 0, g.s0 - 1);
  const int iy = clampi(sat_trunc((y - g.min1) / g.cell1), 0, g.s1 - 1);
  const int iz = clampi(sat_trunc((z - g.min2) / g.cell2), 0, g.s1 - 1);
  return iz * g.s0 * g.s1 + iy * g.s2 + ix;
}

__global__ __launch_bounds__(256) void bin_assign(int n, const double* __restrict__ pos,
                                                  GridArgs g, int* __restrict__ cell) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t N = (size_t)n;
  cell[i] = cell_of(pos[i], pos[N + i], pos[2 * N + i], g);
}

// Per-tile digit histogram, written digit-major: hist[d * ntiles + tile].
__global__ __launch_bounds__(kSortThreads) void radix_hist(int n, const int* __restrict__ keys,
                                                           int stride, int shift, int ntiles,
                                                           int* __restrict__ hist) {
  __shared__ int h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * kSortTile;
  for (int it = 0; it < kSortItems; ++it) {
    const int idx = base + it * kSortThreads + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[(size_t)idx * stride] >> shift) & (kRadix - 1)], 1);
  }
  __syncthreads();
  hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// bin_assign fused with the first pass's tile histogram: each thread loads
// the positions of its kSortItems bodies of the tile before it divides any
// of them, so the loads are in flight together; then the cells go out and
// into the tile's digit histogram (hist[d * ntiles + tile]).
__global__ __launch_bounds__(kSortThreads) void assign_hist(int n, const double* __restrict__ pos,
                                                            GridArgs g, int* __restrict__ cell,
                                                            int shift, int ntiles,
                                                            int* __restrict__ hist) {
  __shared__ int h[kRadix];
  h[threadIdx.x] = 0;
  const size_t N = (size_t)n;
  const int base = blockIdx.x * kSortTile;
  double x[kSortItems], y[kSortItems], z[kSortItems];
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int i = base + it * kSortThreads + threadIdx.x;
    x[it] = i < n ? pos[i] : 0.0;
    y[it] = i < n ? pos[N + i] : 0.0;
    z[it] = i < n ? pos[2 * N + i] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int i = base + it * kSortThreads + threadIdx.x;
    if (i < n) {
      const int c = cell_of(x[it], y[it], z[it], g);
      cell[i] = c;
      atomicAdd(&h[(c >> shift) & (kRadix - 1)], 1);
    }
  }
  __syncthreads();
  hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of 256 values, one per thread of a 256-thread block;
// *total receives the sum.  wsum: 4 ints of LDS.
__device__ __forceinline__ int block_scan256(int v, int* wsum, int* total) {
  const int lane = threadIdx.x & (khrr::kWave - 1);
  const int w = threadIdx.x / khrr::kWave;
  int x = v;
#pragma unroll
  for (int o = 1; o < khrr::kWave; o <<= 1) {
    const int y = __shfl_up(x, o, khrr::kWave);
    if (lane >= o) x += y;
  }
  if (lane == khrr::kWave - 1) wsum[w] = x;
  __syncthreads();
  int before = 0;
  for (int k = 0; k < w; ++k) before += wsum[k];
  *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  __syncthreads();  // wsum is free again for the next call
  return before + x - v;
}

// hist is digit-major, hist[d * ntiles + tile].  Block d scans row d in
// place (exclusive, tile order) and writes the row sum to totals[d]; the
// scatter adds the exclusive prefix of totals, so the whole array is
// scanned in digit-then-tile order with 256 blocks instead of one.
__global__ __launch_bounds__(256) void scan_rows(int ntiles, int* __restrict__ hist,
                                                 int* __restrict__ totals) {
  __shared__ int wsum[4];
  int* row = hist + (size_t)blockIdx.x * ntiles;
  int carry = 0;
  for (int base = 0; base < ntiles; base += 256) {
    const int i = base + (int)threadIdx.x;
    const int v = i < ntiles ? row[i] : 0;
    int total = 0;
    const int ex = block_scan256(v, wsum, &total);
    if (i < ntiles) row[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Stable scatter of one digit pass.  Items are visited in index order
// (round-major, then wave, then lane); a lane's rank among equal digits of
// its wave comes from eight ballots, and the LDS counters carry the running
// per-digit position across waves and rounds.  Every round's key and value
// are loaded before the first round, so the rounds' global loads are in
// flight together instead of one per pair of barriers.  Input: the cells
// with the index as value (pairs_in 0) or interleaved (cell, index) pairs;
// output: interleaved pairs, so each body is one 8-byte store.
__global__ __launch_bounds__(kSortThreads) void radix_scatter(
    int n, const int* __restrict__ keys_in, int pairs_in, int shift,
    int ntiles, const int* __restrict__ offs, const int* __restrict__ totals,
    int2* __restrict__ kv_out) {
  constexpr int kWaves = kSortThreads / khrr::kWave;
  __shared__ int run[kRadix];
  __shared__ int wcnt[kWaves][kRadix];
  __shared__ int wpre[kWaves][kRadix];
  __shared__ int wsum[4];
  const int tid = threadIdx.x;
  const int lane = tid & (khrr::kWave - 1);
  const int wave = tid / khrr::kWave;
  int all = 0;
  const int digit_base = block_scan256(totals[tid], wsum, &all);
  run[tid] = digit_base + offs[(size_t)tid * ntiles + blockIdx.x];
  for (int w = 0; w < kWaves; ++w) wcnt[w][tid] = 0;
  __syncthreads();
  const unsigned long long lt = (1ull << lane) - 1ull;
  const int base = blockIdx.x * kSortTile;
  int keys[kSortItems], vals[kSortItems];
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int idx = base + it * kSortThreads + tid;
    if (pairs_in) {
      const int2 kv = idx < n ? reinterpret_cast<const int2*>(keys_in)[idx] : make_int2(0, 0);
      keys[it] = kv.x;
      vals[it] = kv.y;
    } else {
      keys[it] = idx < n ? keys_in[idx] : 0;
      vals[it] = idx;
    }
  }
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int idx = base + it * kSortThreads + tid;
    const bool valid = idx < n;
    const int key = keys[it];
    const int val = vals[it];
    const int d = (key >> shift) & (kRadix - 1);
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      const bool bit = (d >> b) & 1;
      const unsigned long long bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const int rank = __popcll(peers & lt);
    if (valid && (peers & lt) == 0ull) wcnt[wave][d] = __popcll(peers);
    __syncthreads();
    {
      int b = run[tid];
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        wpre[w][tid] = b;
        b += wcnt[w][tid];
        wcnt[w][tid] = 0;
      }
      run[tid] = b;
    }
    __syncthreads();
    if (valid) {
      const int p = wpre[wave][d] + rank;
      kv_out[p] = make_int2(key, val);
    }
  }
}

__device__ __forceinline__ int lower_bound_dev(const int* __restrict__ a, int st, int n,
                                               long long key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if ((long long)a[(size_t)mid * st] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// First index of a (sorted by cell >> 9) with a[i] >= key: eight-way steps
// (seven independent probes per step, so the chain of dependent loads is
// about a third of a binary search's), then a binary search of the last
// few entries.
__device__ __forceinline__ int lower_bound8(const int* __restrict__ a, int st, int n, long long key) {
  int lo = 0, hi = n;  // the answer lies in [lo, hi]
  while (hi - lo > 16) {
    const int step = (hi - lo) >> 3;  // >= 2: probes lo + step .. lo + 7 step < hi
    int v[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) v[j] = a[(size_t)(lo + (j + 1) * step) * st];
    int below = 0;  // probes < key: a prefix of the seven, as a is sorted
#pragma unroll
    for (int j = 0; j < 7; ++j) below += (long long)v[j] < key ? 1 : 0;
    const int nlo = below == 0 ? lo : lo + below * step + 1;
    const int nhi = below == 7 ? hi : lo + (below + 1) * step;
    lo = nlo;
    hi = nhi;
  }
  return lo + lower_bound_dev(a + (size_t)lo * st, st, hi - lo, key);
}

// bounds[b] = number of sorted particles with cell < b*512, for b in [0, B]:
// all block boundaries found by independent searches in one launch.  The
// keys are every st-th int of skeys (st 2 for interleaved pairs).
__global__ __launch_bounds__(256) void block_bounds(int n, int nb, const int* __restrict__ skeys,
                                                    int st, int* __restrict__ bounds) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b <= nb) bounds[b] = lower_bound8(skeys, st, n, (long long)b * kCellBlock);
}

// One workgroup per 512-cell block of the reference.  The sort grouped the
// bodies by block only, so a block's bodies [lo, hi) arrive in index order
// but not by cell.  skeys holds interleaved (cell, index) pairs (pairs 1) or,
// when no sort ran, the cells in index order (pairs 0).
//  * Empty blocks (most of a fine grid) are four coalesced stores per cell,
//    written non-temporal: they are most of the kernel's traffic (20 B a
//    cell) and would otherwise evict from L2 the sorted keys, indices and
//    masses the occupied blocks read.
//  * An occupied block loads the cell, index and mass of its first kChunk
//    bodies into registers (kReg per thread), so those loads are in flight
//    through the LDS histogram and scan that give counts and exclusive
//    offsets.
//  * Ranks: rounds of 512 bodies in index order.  Within a wave, bodies of
//    the same cell find each other with nine ballots (one per bit of the
//    block-local cell); each group's first lane records the group size in
//    a per-wave byte counter, and a body's rank in its cell is the count of
//    that cell in the rounds before, plus the groups of the waves before
//    it, plus its place in its own group.  That rank is id_in_cell (the
//    index-order rank), and offset + rank its position in cell order.
//  * Mass: each cell adds its bodies' FP64 masses in index order, the sum
//    of a sequential walk.  A block of at most kChunk bodies drops the
//    masses it holds in registers into LDS at their cell-order positions;
//    a larger one writes the cell-order permutation to scratch and sums
//    chunk by chunk.
// Every output cell is written exactly once, so nothing needs clearing.
constexpr int kReg = 4;
constexpr int kCellWaves = kCellBlock / khrr::kWave;

struct CellRanks {
  int* run;                                       // bodies per cell in earlier rounds
  unsigned char (*wc)[kCellWaves][kCellBlock];    // [2] per-wave group sizes
  const int* exs;                                 // block-relative exclusive offsets
};

// One round of ranking: body v (block-local cell k, -1 when the lane has
// none) of this thread, counters of parity `par`.  Every thread of the
// workgroup calls it (two barriers).  Returns the body's position in cell
// order (block-relative); -1 for no body.
__device__ __forceinline__ int rank_round(const CellRanks& cr, int k, int v, int par,
                                          int* __restrict__ id_in_cell) {
  const int tid = threadIdx.x;
  const int lane = tid & (khrr::kWave - 1);
  const int wave = tid / khrr::kWave;
  const bool valid = k >= 0;
  unsigned long long peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 9; ++b) {  // 512 cells per block: nine bits
    const bool bit = (k >> b) & 1;
    const unsigned long long bal = __ballot(bit);
    peers &= bit ? bal : ~bal;
  }
  const unsigned long long lt = (1ull << lane) - 1ull;
  if (valid && (peers & lt) == 0ull) cr.wc[par][wave][k] = (unsigned char)__popcll(peers);
  __syncthreads();
  int sp = -1;
  if (valid) {
    int rank = cr.run[k] + __popcll(peers & lt);
    for (int w = 0; w < wave; ++w) rank += cr.wc[par][w][k];
    if (id_in_cell) id_in_cell[v] = rank;
    sp = cr.exs[k] + rank;
  }
  __syncthreads();
  // thread t folds cell t's groups of this round into its running count
  // and clears them; the round after next writes this parity again
  int add = 0;
#pragma unroll
  for (int w = 0; w < kCellWaves; ++w) {
    add += cr.wc[par][w][tid];
    cr.wc[par][w][tid] = 0;
  }
  cr.run[tid] += add;
  return sp;
}

__global__ __launch_bounds__(kCellBlock) void bin_cells(
    long long G, const int* __restrict__ skeys, int pairs,
    const int* __restrict__ bounds, const double* __restrict__ mass, int* __restrict__ count,
    int* __restrict__ offset, int* __restrict__ start_ref, int* __restrict__ block_incl,
    int* __restrict__ id_in_cell, double* __restrict__ cell_mass, int* __restrict__ perm) {
  constexpr int kChunk = kReg * kCellBlock;
  __shared__ int cnt[kCellBlock];
  __shared__ int exs[kCellBlock];
  __shared__ int run[kCellBlock];
  __shared__ unsigned char wc[2][kCellWaves][kCellBlock];
  __shared__ int wsum[kCellWaves];
  __shared__ double msh[kChunk];
  const int tid = threadIdx.x;
  const int blk = blockIdx.x;
  const long long c0 = (long long)blk * kCellBlock;
  const long long c = c0 + tid;
  const int lo = bounds[blk], hi = bounds[blk + 1];
  if (tid == 0 && block_incl) block_incl[blk] = hi;
  if (lo == hi) {
    if (c < G) {
      __builtin_nontemporal_store(0, &count[c]);
      if (offset) __builtin_nontemporal_store(lo, &offset[c]);
      if (start_ref) __builtin_nontemporal_store(lo, &start_ref[c]);
      if (cell_mass) __builtin_nontemporal_store(0.0, &cell_mass[c]);
    }
    return;
  }
  const int nb = hi - lo;
  const bool small = nb <= kChunk;  // the same for the whole workgroup
  // the first chunk's cells (relative to the block), indices and masses
  int key[kReg], val[kReg];
  double mv[kReg];
#pragma unroll
  for (int u = 0; u < kReg; ++u) {
    const int p = lo + u * kCellBlock + tid;
    if (pairs) {
      const int2 kv = p < hi ? reinterpret_cast<const int2*>(skeys)[p] : make_int2(0, 0);
      key[u] = p < hi ? kv.x - (int)c0 : -1;
      val[u] = kv.y;
    } else {
      key[u] = p < hi ? skeys[p] - (int)c0 : -1;
      val[u] = p < hi ? p : 0;
    }
  }
  if (cell_mass && small) {
#pragma unroll
    for (int u = 0; u < kReg; ++u) mv[u] = key[u] >= 0 ? mass[val[u]] : 0.0;
  }
  cnt[tid] = 0;
  run[tid] = 0;
#pragma unroll
  for (int w = 0; w < kCellWaves; ++w) wc[0][w][tid] = wc[1][w][tid] = 0;
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kReg; ++u)
    if (key[u] >= 0) atomicAdd(&cnt[key[u]], 1);
  for (int p = lo + kChunk + tid; p < hi; p += kCellBlock)
    atomicAdd(&cnt[skeys[(size_t)p << pairs] - (int)c0], 1);
  __syncthreads();
  // Exclusive scan of cnt: wave-inclusive scans, then wave totals.
  const int lane = tid & (khrr::kWave - 1);
  const int wave = tid / khrr::kWave;
  const int mine = cnt[tid];
  int incl = mine;
#pragma unroll
  for (int off = 1; off < khrr::kWave; off <<= 1) {
    const int v = __shfl_up(incl, off, khrr::kWave);
    if (lane >= off) incl += v;
  }
  if (lane == khrr::kWave - 1) wsum[wave] = incl;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += wsum[w];
  const int excl = before + incl - mine;  // block-relative
  exs[tid] = excl;
  if (c < G) {
    count[c] = mine;
    if (offset) offset[c] = lo + excl;
    if (start_ref) start_ref[c] = lo;
  }
  // (the first rank round's first barrier publishes exs)
  const CellRanks cr{run, wc, exs};
  const bool want_perm = cell_mass && !small;
#pragma unroll
  for (int u = 0; u < kReg; ++u) {
    if (u * kCellBlock >= nb) break;  // uniform: every thread leaves together
    const int sp = rank_round(cr, key[u], val[u], u & 1, id_in_cell);
    if (sp >= 0) {
      if (cell_mass && small) msh[sp] = mv[u];
      if (want_perm) perm[lo + sp] = val[u];
    }
  }
  // rounds past the first chunk read their bodies from the sorted keys;
  // the next round's loads are issued before this round's barriers
  if (lo + kChunk < hi) {
    int p = lo + kChunk + tid;
    int k = p < hi ? skeys[(size_t)p << pairs] - (int)c0 : -1;
    int v = p < hi ? (pairs ? skeys[((size_t)p << 1) + 1] : p) : 0;
    for (int r = kReg, p0 = lo + kChunk; p0 < hi; ++r, p0 += kCellBlock) {
      const int pn = p0 + kCellBlock + tid;
      const int kn = pn < hi ? skeys[(size_t)pn << pairs] - (int)c0 : -1;
      const int vn = pn < hi ? (pairs ? skeys[((size_t)pn << 1) + 1] : pn) : 0;
      const int sp = rank_round(cr, k, v, r & 1, id_in_cell);
      if (sp >= 0 && want_perm) perm[lo + sp] = v;
      k = kn;
      v = vn;
    }
  }
  if (!cell_mass) return;
  __syncthreads();  // msh / perm complete
  double m = 0.0;
  if (small) {
    for (int q = excl; q < excl + mine; ++q) m += msh[q];
  } else {
    // cell order in chunks of kChunk positions: gather the chunk's masses
    // into LDS, then each cell adds its part of the chunk in order
    for (int q0 = 0; q0 < nb; q0 += kChunk) {
      double mm[kReg];
#pragma unroll
      for (int u = 0; u < kReg; ++u) {
        const int q = q0 + u * kCellBlock + tid;
        mm[u] = q < nb ? mass[perm[lo + q]] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kReg; ++u) msh[u * kCellBlock + tid] = mm[u];
      __syncthreads();
      const int q1 = min(excl + mine, q0 + kChunk);
      for (int q = max(excl, q0); q < q1; ++q) m += msh[q - q0];
      __syncthreads();  // msh is refilled by the next chunk
    }
  }
  if (c < G) cell_mass[c] = m;
}

// perm[offset[cell_id[i]] + id_in_cell[i]] = i: the sorted segments back
// from the binning outputs.  Entries that do not come from a consistent
// binning (index outside the grid or the particle range) are skipped.
__global__ __launch_bounds__(256) void segments_from_bins(int n, long long G,
                                                          const int* __restrict__ cell_id,
                                                          const int* __restrict__ id_in_cell,
                                                          const int* __restrict__ offset,
                                                          int* __restrict__ perm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cell_id[i];
  if (c < 0 || c >= G) return;
  const long long p = (long long)offset[c] + id_in_cell[i];
  if (p >= 0 && p < n) perm[p] = i;
}

// A cell's walk over its sorted segment [o, o + k) stops where the segment
// leaves [0, n) (only an inconsistent binning does that); this is the
// number of entries it visits.
__device__ __forceinline__ int segment_len(long long o, int k, int n) {
  if (o < 0 || o >= n) return 0;
  return (int)min((long long)k, (long long)n - o);
}
constexpr int kWalk = 4;

// computeCellPhiUnsorted (p2mesh-kernel.cuh:152-167) in index order.
__global__ __launch_bounds__(256) void cell_phi_kernel(int n, long long G,
                                                       const int* __restrict__ count,
                                                       const int* __restrict__ offset,
                                                       const int* __restrict__ perm,
                                                       const float* __restrict__ phi,
                                                       double* __restrict__ cell_phi,
                                                       int accumulate) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G) return;
  const int k = count[c];
  double v = accumulate ? cell_phi[c] : 0.0;
  if (k > 0) {
    const long long o = offset[c];
    const int kk = segment_len(o, k, n);
    int q = 0;
    for (; q + kWalk <= kk; q += kWalk) {  // kWalk loads in flight, adds in order
      int i[kWalk];
      float x[kWalk];
#pragma unroll
      for (int u = 0; u < kWalk; ++u) i[u] = perm[o + q + u];
#pragma unroll
      for (int u = 0; u < kWalk; ++u) x[u] = (i[u] >= 0 && i[u] < n) ? phi[i[u]] : 0.f;
#pragma unroll
      for (int u = 0; u < kWalk; ++u)
        if (i[u] >= 0 && i[u] < n) v += (double)x[u] / k;
    }
    for (; q < kk; ++q) {
      const int i = perm[o + q];
      if (i >= 0 && i < n) v += (double)phi[i] / k;
    }
  }
  // non-temporal, as the per-cell outputs of the binning: most cells are
  // empty, and this stream would evict the segments the others read
  __builtin_nontemporal_store(v, &cell_phi[c]);
}

// apply_particle_to_cell_weighted (p2mesh-kernel.cuh:273-287) in index
// order: sum of value[i] * (mass[i] / cell_mass[c]) over the cell's bodies.
__global__ __launch_bounds__(256) void cell_weighted_kernel(int n, long long G,
                                                            const int* __restrict__ count,
                                                            const int* __restrict__ offset,
                                                            const int* __restrict__ perm,
                                                            const double* __restrict__ cell_mass,
                                                            const double* __restrict__ mass,
                                                            const double* __restrict__ value,
                                                            double* __restrict__ cell_value) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G) return;
  const int k = count[c];
  double v = 0.0;
  if (k > 0) {
    const double cm = cell_mass[c];
    const long long o = offset[c];
    const int kk = segment_len(o, k, n);
    int q = 0;
    for (; q + kWalk <= kk; q += kWalk) {  // kWalk loads in flight, adds in order
      int i[kWalk];
      double x[kWalk], m[kWalk];
#pragma unroll
      for (int u = 0; u < kWalk; ++u) i[u] = perm[o + q + u];
#pragma unroll
      for (int u = 0; u < kWalk; ++u) {
        const bool ok = i[u] >= 0 && i[u] < n;
        x[u] = ok ? value[i[u]] : 0.0;
        m[u] = ok ? mass[i[u]] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kWalk; ++u)
        if (i[u] >= 0 && i[u] < n) v += x[u] * (m[u] / cm);
    }
    for (; q < kk; ++q) {
      const int i = perm[o + q];
      if (i >= 0 && i < n) v += value[i] * (mass[i] / cm);
    }
  }
  __builtin_nontemporal_store(v, &cell_value[c]);
}

// apply_cell_to_particles_avg (p2mesh-kernel.cuh:256-271): the cell's value
// shared out over its bodies, -1 for a cell of no bodies, NaN for an index
// outside the grid.
__global__ __launch_bounds__(256) void cell_avg_kernel(int n, long long G,
                                                       const int* __restrict__ cell_id,
                                                       const int* __restrict__ count,
                                                       const double* __restrict__ cell_value,
                                                       double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cell_id[i];
  if (c < 0 || c >= G) {
    out[i] = __builtin_nan("");
    return;
  }
  const int k = count[c];
  out[i] = k > 0 ? cell_value[c] / k : -1.0;
}

}  // namespace

extern "C" int khrr_p2mesh_cell_phi(khrr_ctx* ctx, int n, long long n_cells, const int* cell_id,
                                    const int* id_in_cell, const int* count, const int* offset,
                                    const float* phi, double* cell_phi, int accumulate) {
  using namespace khrr;
  KHRR_TRY(enter(ctx));
  if (n < 0 || n_cells < 1 || n_cells > INT_MAX)
    return fail(ctx, KHRR_EINVAL, "bad sizes n=%d cells=%lld", n, n_cells);
  if (!count || !offset || !cell_phi || (n > 0 && (!cell_id || !id_in_cell || !phi)))
    return fail(ctx, KHRR_EINVAL, "null pointer");
  void* base = nullptr;
  KHRR_TRY(scratch(ctx, Carver::need((size_t)(n > 0 ? n : 1), sizeof(int)), &base));
  int* perm = static_cast<int*>(base);
  Timed tm(ctx, KHRR_K_BIN);
  if (n > 0)
    hipLaunchKernelGGL(segments_from_bins, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, n,
                       n_cells, cell_id, id_in_cell, offset, perm);
  hipLaunchKernelGGL(cell_phi_kernel, dim3(ceil_div(n_cells, 256)), dim3(256), 0, ctx->stream, n,
                     n_cells, count, offset, perm, phi, cell_phi, accumulate);
  return check_launch(ctx, "cell_phi kernels");
}

extern "C" int khrr_p2mesh_cell_weighted(khrr_ctx* ctx, int n, long long n_cells,
                                         const int* cell_id, const int* id_in_cell,
                                         const int* count, const int* offset,
                                         const double* cell_mass, const double* mass,
                                         const double* value, double* cell_value) {
  using namespace khrr;
  KHRR_TRY(enter(ctx));
  if (n < 0 || n_cells < 1 || n_cells > INT_MAX)
    return fail(ctx, KHRR_EINVAL, "bad sizes n=%d cells=%lld", n, n_cells);
  if (!count || !offset || !cell_mass || !cell_value ||
      (n > 0 && (!cell_id || !id_in_cell || !mass || !value)))
    return fail(ctx, KHRR_EINVAL, "null pointer");
  void* base = nullptr;
  KHRR_TRY(scratch(ctx, Carver::need((size_t)(n > 0 ? n : 1), sizeof(int)), &base));
  int* perm = static_cast<int*>(base);
  Timed tm(ctx, KHRR_K_BIN);
  if (n > 0)
    hipLaunchKernelGGL(segments_from_bins, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, n,
                       n_cells, cell_id, id_in_cell, offset, perm);
  hipLaunchKernelGGL(cell_weighted_kernel, dim3(ceil_div(n_cells, 256)), dim3(256), 0,
                     ctx->stream, n, n_cells, count, offset, perm, cell_mass, mass, value,
                     cell_value);
  return check_launch(ctx, "cell_weighted kernels");
}

extern "C" int khrr_p2mesh_cell_avg(khrr_ctx* ctx, int n, long long n_cells, const int* cell_id,
                                    const int* count, const double* cell_value, double* out) {
  using namespace khrr;
  KHRR_TRY(enter(ctx));
  if (n < 0 || n_cells < 1 || n_cells > INT_MAX)
    return fail(ctx, KHRR_EINVAL, "bad sizes n=%d cells=%lld", n, n_cells);
  if (n > 0 && (!cell_id || !count || !cell_value || !out))
    return fail(ctx, KHRR_EINVAL, "null pointer");
  if (n == 0) return 0;
  Timed tm(ctx, KHRR_K_BIN);
  hipLaunchKernelGGL(cell_avg_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, n,
                     n_cells, cell_id, count, cell_value, out);
  return check_launch(ctx, "cell_avg_kernel");
}

namespace khrr {

int check_grid(khrr_ctx* ctx, const khrr_grid* grid) {
  if (!grid) return fail(ctx, KHRR_EINVAL, "grid is NULL");
  const long long G = (long long)grid->size[0] * grid->size[1] * grid->size[2];
  if (grid->size[0] <= 0 || grid->size[1] <= 0 || grid->size[2] <= 0 || G > INT_MAX)
    return fail(ctx, KHRR_EINVAL, "invalid grid size %d x %d x %d", grid->size[0], grid->size[1],
                grid->size[2]);
  // The reference clamps z with size[1] and strides y by size[2]; for a
  // non-cubic grid that can index past G, which is rejected here.
  const long long max_cell = (long long)(grid->size[1] - 1) * grid->size[0] * grid->size[1] +
                             (long long)(grid->size[1] - 1) * grid->size[2] + (grid->size[0] - 1);
  if (max_cell >= G)
    return fail(ctx, KHRR_EINVAL,
                "grid %d x %d x %d: the reference's cell index (z clamped by size[1], y "
                "stride size[2]) can exceed the cell count for this shape",
                grid->size[0], grid->size[1], grid->size[2]);
  return 0;
}

}  // namespace khrr

namespace khrr {

int bin_validate(khrr_ctx* ctx, const khrr_grid* grid, int n, const double* pos,
                 const double* mass, const int* count, const double* cell_mass) {
  KHRR_TRY(check_grid(ctx, grid));
  if (n < 0) return fail(ctx, KHRR_EINVAL, "n < 0");
  if (!count) return fail(ctx, KHRR_EINVAL, "count is NULL");
  if (n > 0 && !pos) return fail(ctx, KHRR_EINVAL, "pos is NULL");
  if (n > 0 && cell_mass && !mass) return fail(ctx, KHRR_EINVAL, "mass is NULL");
  return 0;
}

int bin_launch(khrr_ctx* ctx, const khrr_grid* grid, int n, const double* pos, const double* mass,
               int* cell_id, int* id_in_cell, int* count, int* offset, int* start_id_ref,
               int* block_incl, double* cell_mass, void* base, bool timed) {
  static_assert(kBinSortTile == kSortTile && kBinRadix == kRadix, "scratch layout constants");
  const long long G = (long long)grid->size[0] * grid->size[1] * grid->size[2];
  const int ntiles = n > 0 ? (int)ceil_div(n, kSortTile) : 1;
  const int cblocks = (int)ceil_div(G, kCellBlock);
  Carver cv(base);
  int* cells = cv.take<int>(n > 0 ? n : 1);
  int2* kva = cv.take<int2>(n > 0 ? n : 1);
  int2* kvb = cv.take<int2>(n > 0 ? n : 1);
  int* hist = cv.take<int>((size_t)kRadix * ntiles);
  int* totals = cv.take<int>((size_t)kRadix);
  int* bounds = cv.take<int>((size_t)cblocks + 1);
  if (cv.used > bin_scratch_need(n, G))
    return fail(ctx, KHRR_EINVAL, "binning scratch layout exceeds its size");
  if (cell_id) cells = cell_id;

  GridArgs g{grid->min[0], grid->min[1], grid->min[2], grid->cell[0], grid->cell[1],
             grid->cell[2], grid->size[0], grid->size[1], grid->size[2]};
  // The sort orders bodies by 512-cell block only (stable, so a block's
  // bodies stay in index order): digits of cell >> 9.  bin_cells orders
  // each block by cell.
  int bits = 0;
  while (bits < 31 && (1ll << bits) < cblocks) ++bits;
  const int passes = (bits + kRadixBits - 1) / kRadixBits;

  Timed tm(timed ? ctx : nullptr, KHRR_K_BIN);
  const int* skeys = cells;
  int pairs = 0;                                 // plain keys, identity values
  int* perm = reinterpret_cast<int*>(kva);       // the pair buffer the sort leaves free
  if (n > 0) {
    if (passes == 0)
      hipLaunchKernelGGL(bin_assign, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, n, pos, g,
                         cells);
    const int* kin = cells;
    for (int p = 0; p < passes; ++p) {
      int2* kvout = (p & 1) ? kvb : kva;
      const int shift = kBlockShift + p * kRadixBits;
      if (p == 0)  // the cells and the first pass's tile histogram in one launch
        hipLaunchKernelGGL(assign_hist, dim3(ntiles), dim3(kSortThreads), 0, ctx->stream, n, pos,
                           g, cells, shift, ntiles, hist);
      else
        hipLaunchKernelGGL(radix_hist, dim3(ntiles), dim3(kSortThreads), 0, ctx->stream, n, kin,
                           2, shift, ntiles, hist);
      hipLaunchKernelGGL(scan_rows, dim3(kRadix), dim3(256), 0, ctx->stream, ntiles, hist,
                         totals);
      hipLaunchKernelGGL(radix_scatter, dim3(ntiles), dim3(kSortThreads), 0, ctx->stream, n, kin,
                         p > 0 ? 1 : 0, shift, ntiles, hist, totals, kvout);
      kin = reinterpret_cast<const int*>(kvout);
      perm = reinterpret_cast<int*>((p & 1) ? kva : kvb);
    }
    skeys = kin;
    pairs = passes > 0 ? 1 : 0;
  }
  hipLaunchKernelGGL(block_bounds, dim3(ceil_div(cblocks + 1, 256)), dim3(256), 0, ctx->stream, n,
                     cblocks, skeys, pairs ? 2 : 1, bounds);
  hipLaunchKernelGGL(bin_cells, dim3(cblocks), dim3(kCellBlock), 0, ctx->stream, G, skeys, pairs,
                     bounds, mass, count, offset, start_id_ref, block_incl, id_in_cell,
                     cell_mass, perm);
  return check_launch(ctx, "p2mesh kernels");
}

}  // namespace khrr

extern "C" int khrr_p2mesh_bin(khrr_ctx* ctx, const khrr_grid* grid, int n, const double* pos,
                               const double* mass, int* cell_id, int* id_in_cell, int* count,
                               int* offset, int* start_id_ref, int* block_incl,
                               double* cell_mass) {
  using namespace khrr;
  KHRR_TRY(enter(ctx));
  KHRR_TRY(bin_validate(ctx, grid, n, pos, mass, count, cell_mass));
  const long long G = (long long)grid->size[0] * grid->size[1] * grid->size[2];
  void* base = nullptr;
  KHRR_TRY(scratch(ctx, bin_scratch_need(n, G), &base));
  return bin_launch(ctx, grid, n, pos, mass, cell_id, id_in_cell, count, offset, start_id_ref,
                    block_incl, cell_mass, base, true);
}
