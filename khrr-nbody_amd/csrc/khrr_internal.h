// khrr_internal.h -- shared plumbing of the gfx950 library: the context
// object (stream, error text, scratch arena, kernel timers, tuning cache)
// and small helpers used by every kernel file.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "khrr.h"

namespace khrr {

constexpr int kWave = 64;  // CDNA wavefront width

// Pair-symmetric form: one (A, B) superblock item of the system
// whose bodies start at `base`; its superblock flags start at
// flags[flag_base] and its partial slots at part[part] (n_sb*4*n floats).
struct SymItem {
  int base, n, A, B, flag_base, pad;
  long long part;
};
// One system of a many-system integrator launch: FP64 state planes pos,
// vel, pos_t, vel_t (3n each) from `state`, FP32 planes at offset `begin`,
// blocks [first_block, first_block + ceil(n/256)).
struct IntSeg {
  double* state;
  double dt;
  int n, begin, first_block, pad;
};
// A system of the pair-symmetric batched form, for the slot reduction.
struct SymSys {
  int base, n, n_sb, pad;
  long long part;
};

struct TimerPair {
  hipEvent_t start;
  hipEvent_t stop;
  int cls;
};

// A captured batched gravity launch sequence (KHRR_GRAPH): its key (the
// composition, every pointer and shape it was captured with, and the
// scratch arena it uses) and a private copy of its device work list.
struct GraphEntry {
  std::vector<long long> key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void* dev = nullptr;  // work list copy (items, ranges, flags)
  unsigned long long last_use = 0;
  bool failed = false;  // capture not possible: this key runs eagerly
};

}  // namespace khrr

struct khrr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int cu_count = 256;
  std::string arch;  // gcnArchName of the device (keys the tuning cache file)
  // Scratch arena: one device allocation grown on demand; callers carve
  // stream-ordered sub-buffers out of it for the duration of one call.
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // Small persistent device words (flags, counters): 256 B.
  int* dev_words = nullptr;
  // Pinned host staging for synchronous results.
  double* host_stage = nullptr;
  // Kernel timing.
  bool timing = false;
  std::vector<khrr::TimerPair> pending;
  std::vector<hipEvent_t> event_pool;
  long launches[KHRR_K_COUNT] = {0};
  double total_ms[KHRR_K_COUNT] = {0};
  // Autotuner cache, indexed by ceil(log2(n)) (kTuneBuckets entries).
  khrr_tune tuned[40] = {};
  bool has_tuned[40] = {};
  double tuned_ms[40] = {};
  khrr_tune batch_tuned = {};
  bool has_batch_tuned = false;
  // Batched form: the work list (system, i-group) of the last composition,
  // kept on the device until the composition or the i-group width changes.
  std::vector<int> batch_off;
  int batch_ipl = 0;
  int batch_items = 0;
  int batch_tiles = 0;
  int batch_red = 0;          // systems whose j-range is cut into pieces
  int batch_red_max = 0;      // largest such system
  long long batch_units = 0;  // partial-sum rows of all cut systems
  int batch_sym_items = 0;    // pair-symmetric (A, B) items
  int batch_sym_big = 0;      // of which the first ones have 4096-body superblocks,
  int batch_sym_mid = 0;      // the next ones 2048 (the rest 1024)
  int batch_sym_sys = 0;      // systems on the pair-symmetric path
  int batch_sym_max = 0;      // largest such system
  long long batch_sym_floats = 0;  // their partial slots
  void* batch_dev = nullptr;  // items, j-tile ranges, cut systems, tile flags
  size_t batch_dev_bytes = 0;
  void* batch_pinned = nullptr;
  size_t batch_pinned_bytes = 0;
  // Captured batched launches (KHRR_GRAPH), least recently used evicted.
  std::vector<khrr::GraphEntry> graphs;
  unsigned long long graph_clock = 0;
  hipStream_t capture_stream = nullptr;  // private, non-blocking, default priority
  // Second stream for concurrent parts of one call (forked from and joined
  // back to `stream` with the two events); non-blocking, default priority.
  hipStream_t aux_stream = nullptr;
  hipEvent_t fork_event = nullptr;
  hipEvent_t join_event = nullptr;
};

namespace khrr {

// Records an error on ctx (or the thread-local slot when ctx is null) and
// returns `code`.
int fail(khrr_ctx* ctx, int code, const char* fmt, ...);
// Thread-local message for calls without a context.
std::string& thread_error();

// Makes the context's device current for the calling host thread.
int enter(khrr_ctx* ctx);

// Ensures the scratch arena holds at least `bytes`; returns its base.
int scratch(khrr_ctx* ctx, size_t bytes, void** base);

// Creates the context's second stream and its fork/join events on first use.
int aux_stream(khrr_ctx* ctx);

// Bump allocator over a scratch region (256-byte aligned pieces).
struct Carver {
  char* base;
  size_t used = 0;
  explicit Carver(void* b) : base(static_cast<char*>(b)) {}
  template <typename T>
  T* take(size_t count) {
    T* p = reinterpret_cast<T*>(base + used);
    used += (count * sizeof(T) + 255) & ~size_t(255);
    return p;
  }
  static size_t need(size_t count, size_t elem) { return (count * elem + 255) & ~size_t(255); }
};

// RAII bracket that records start/stop events around the launches made in
// its scope when the context has timing enabled.
class Timed {
 public:
  Timed(khrr_ctx* ctx, int cls);
  ~Timed();

 private:
  khrr_ctx* ctx_;
  int cls_;
  hipEvent_t start_ = nullptr;
};

// Checks the launch that was just made.
int check_launch(khrr_ctx* ctx, const char* what);

inline unsigned ceil_div(long long a, long long b) { return unsigned((a + b - 1) / b); }

// Gravity launch helpers shared by the plain entry point, the batched form
// and the autotuner (khrr_gravity.hip).
int validate_tune(khrr_ctx* ctx, const khrr_tune& t);
khrr_tune default_tune(const khrr_ctx* ctx, int n_i, int n_j);
const khrr_tune* cached_tune(const khrr_ctx* ctx, int n);
constexpr int kTuneBuckets = 40;
// Size bucket of the tuning cache: ceil(log2(n)), at most kTuneBuckets - 1.
int tune_bucket(int n);
void store_tune(khrr_ctx* ctx, int n, const khrr_tune& t);
int accel_launch(khrr_ctx* ctx, int n_i, const float* pos_i, const float* h_i, int n_j,
                 const float* pos_j, const float* m_j, const float* h_j, float* acc, float* phi,
                 unsigned flags, const khrr_tune& tune);
// Predictor/corrector (khrr_integrate.hip); ld32 is the plane stride of the
// FP32 arrays acc, acc0 and pos32 (n for the public entry point).
int integrate_launch(khrr_ctx* ctx, int n, int istep, double dt, const khrr_extpot* ext,
                     const float* acc, double* pos_t, double* vel_t, const double* pos,
                     const double* vel, float* acc0, float* pos32, int ld32);
// Predictor (istep 0) or corrector (istep 1) of many systems in one launch
// over a device segment table (no external potential).
int integrate_segments(khrr_ctx* ctx, int n_seg, const IntSeg* segs, int total_blocks, int istep,
                       const float* acc, float* acc0, float* pos32, int ld32);
// Batched form over arbitrary segments [begin, begin+count) of planes with
// stride ld (khrr_batch.hip); host arrays describe the segments.
int batched_segments(khrr_ctx* ctx, int n_sys, const int* host_begin, const int* host_count,
                     int ld, const float* pos, const float* h, const float* m, float* acc,
                     float* phi, unsigned flags, const khrr_tune& t);
// i-bodies per lane for one batched evaluation of these systems: the
// widest of pref, pref/2, ..., 1 whose one-sided items number at least two
// per SIMD, else 1, so a round of a few small systems lasts about one
// short item.  Results do not depend on it (each body's j order is fixed
// by its system's j-pieces).
int round_ipl(const khrr_ctx* ctx, int n_sys, const int* count, int pref);
// Scratch bytes the batched form needs for the partial sums of one system
// of n bodies (0 when it writes its outputs directly).
size_t batched_part_bytes(int n, bool compensated);
// Binning (khrr_p2mesh.hip).  Scratch of one call of n bodies on G cells,
// in 256-byte aligned pieces: the cell indices, the radix sort's two key
// and two value buffers, its tile histograms and digit totals, and the
// bounds of the 512-cell blocks.  constexpr, so a fixed buffer can be sized
// from it.
constexpr int kBinSortTile = 2048;  // bodies per sort tile
constexpr int kBinRadix = 256;      // digits per sort pass
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t bin_scratch_need(long long n, long long G) {
  const long long nn = n > 0 ? n : 1;
  const long long ntiles = n > 0 ? (n + kBinSortTile - 1) / kBinSortTile : 1;
  const long long cblocks = (G + KHRR_BIN_BLOCK - 1) / KHRR_BIN_BLOCK;
  return 5 * align256((size_t)nn * 4) + align256((size_t)kBinRadix * ntiles * 4) +
         align256((size_t)kBinRadix * 4) + align256((size_t)(cblocks + 1) * 4);
}
// Argument checks of khrr_p2mesh_bin (the grid via check_grid).
int bin_validate(khrr_ctx* ctx, const khrr_grid* grid, int n, const double* pos,
                 const double* mass, const int* count, const double* cell_mass);
// Enqueues one binning call with its scratch at `base` (at least
// bin_scratch_need(n, G) bytes, 256-byte aligned).  timed: bracket it with
// the context's kernel timers (their events come from a pool that may
// create one).
int bin_launch(khrr_ctx* ctx, const khrr_grid* grid, int n, const double* pos, const double* mass,
               int* cell_id, int* id_in_cell, int* count, int* offset, int* start_id_ref,
               int* block_incl, double* cell_mass, void* base, bool timed);
// Looks up, once per device, the device address of the fixed scratch of
// khrr_p2mesh_bin_noalloc (khrr_noalloc.hip); called when a context is
// created, with its device current.
int noalloc_init(khrr_ctx* ctx);
// Grid shapes khrr_p2mesh_bin accepts (khrr_p2mesh.hip): every size >= 1,
// at most INT_MAX cells, and the reference's cell index (z clamped by
// size[1], y stride size[2]) never past the last cell.
int check_grid(khrr_ctx* ctx, const khrr_grid* grid);
// Pair-symmetric form (khrr_sym.hip) for one system (i-set == j-set).
// Bodies per superblock: `requested` when it is 1024, 2048 or 4096, else
// the rule of n alone -- 4096 from 524288 bodies on, 2048 from 32768, 1024
// below (shared by the plain and batched forms; the batched form always
// follows the rule).
int sym_superblock(int n, int requested);
// i-bodies per lane the pair-symmetric kernel runs for superblocks of sb
// bodies when asked for ipl (1024-body superblocks: always 4).
int sym_ipl(int sb, int ipl);
bool sym_eligible(int n, int ipl, int requested);
bool sym_worthwhile(const khrr_ctx* ctx, int n);
int sym_launch(khrr_ctx* ctx, int n, const float* pos, const float* h, const float* m, float* acc,
               float* phi, unsigned flags, int ipl, int requested);
// Batched pair-symmetric launches (i_per_lane 8) over a device item table
// holding n_big items of 4096-body superblocks then n_small of 2048, then
// the per-system slot reduction; max_n bounds every system.
int sym_items_launch(khrr_ctx* ctx, int n_big, int n_mid, int n_small, const SymItem* items,
                     int n_sys, const SymSys* sys, int max_n, int ld, const float* pos,
                     const float* h, const float* m, const int* sb_flags, float* part, float* acc,
                     float* phi, unsigned flags);

}  // namespace khrr

#define KHRR_TRY(expr)       \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != 0) return rc_; \
  } while (0)

#define KHRR_HIP(ctx, call)                                                              \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return khrr::fail((ctx), KHRR_EDEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)
