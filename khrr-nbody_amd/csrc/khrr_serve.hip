// khrr_serve.hip -- serving form: continuous batching of many simulations.
//
// The reference runs one simulation per process (nbody.cu:1010-1711).  A
// server here owns a device arena for the FP32 gravity inputs and outputs
// of every admitted simulation (48 B per body: positions, mass, softening,
// two acceleration buffers, potential) plus each simulation's FP64 state
// (104 B per body), all held within a byte budget.  Each khrr_server_step
// is one scheduling round:
//   1. admission, first come first served: the head request is admitted
//      when its footprint fits the budget, a contiguous arena range is free
//      and an active slot is free; new systems get their initial force,
//      potential and t = 0 diagnostics;
//   2. one step for the round's systems: every active system, or (with a
//      round_pairs target) those with the least remaining work until the
//      round holds that many pair-interactions; ONE predictor launch, ONE
//      batched gravity evaluation (replayed as a captured HIP graph while
//      the round's composition repeats) and ONE corrector launch;
//   3. systems that reached their step count get their potential, final
//      diagnostics and the p2mesh mass grid, and release their memory.
// Only rounds that admit or finish wait for the device; the others return
// once queued (two segment tables in a ring bound the run-ahead to two
// rounds), so host scheduling overlaps the device's work.
// A system's arithmetic never mixes with another's (segment-local j loops,
// per-system flags, per-system reductions), so its results are
// bit-identical to running it alone.
#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "khrr_internal.h"

namespace {

constexpr size_t kArenaFloatsPerBody = 12;  // pos32 3, m32, h32, acc 3, acct 3, phi
constexpr size_t kStateDoublesPerBody = 13; // pos 3, vel 3, post 3, velt 3, mass
constexpr size_t kBytesPerBody = kArenaFloatsPerBody * 4 + kStateDoublesPerBody * 8;  // 152

// Device bytes a simulation of n bodies holds while active: arena share,
// FP64 state and the batched gravity's partial sums for it.
size_t footprint(int n) { return kBytesPerBody * (size_t)n + khrr::batched_part_bytes(n, false); }

enum class State { kPending, kActive, kFinished };

struct Sim {
  long long id = 0;
  int n = 0;
  double dt = 0.0;
  int steps = 0;
  khrr_grid grid{};
  State state = State::kPending;
  std::vector<double> pos, vel, mass;
  std::vector<float> m32, h32, pos32;
  int begin = -1;
  double* dev = nullptr;  // FP64 planes: pos, vel, post, velt, then mass
  int done_steps = 0;
  double diag0[KHRR_DIAG_COUNT] = {};
  double diag1[KHRR_DIAG_COUNT] = {};
  std::vector<double> cell_mass;
  double* d_pos() const { return dev; }
  double* d_vel() const { return dev + 3 * (size_t)n; }
  double* d_post() const { return dev + 6 * (size_t)n; }
  double* d_velt() const { return dev + 9 * (size_t)n; }
  double* d_mass() const { return dev + 12 * (size_t)n; }
};

// First-fit allocator of body ranges in [0, cap).
class Ranges {
 public:
  void init(int cap) {
    free_.clear();
    if (cap > 0) free_[0] = cap;
  }
  int take(int n) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const int b = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[b + n] = len - n;
        return b;
      }
    }
    return -1;
  }
  void give(int b, int n) {
    if (n <= 0) return;
    auto it = free_.emplace(b, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<int, int> free_;
};

}  // namespace

struct khrr_server {
  khrr_ctx* ctx = nullptr;
  size_t budget = 0;
  int max_active = 1;
  int cap = 0;
  float* arena = nullptr;
  float *pos32 = nullptr, *m32 = nullptr, *h32 = nullptr, *acc = nullptr, *acct = nullptr,
        *phi = nullptr;
  Ranges ranges;
  size_t in_use = 0;
  std::deque<Sim*> pending;
  std::vector<Sim*> active;
  std::map<long long, std::unique_ptr<Sim>> sims;
  long long next_id = 1;
  long long rounds = 0;
  double pairs = 0.0;
  int* bin_count = nullptr;
  double* bin_mass = nullptr;
  size_t bin_cells = 0;
  khrr_tune tune{4, 4, 1, 256, 0};
  // Round scheduling: least remaining work first until the round holds
  // round_pairs pair-interactions per step (0: every active system).
  double round_pairs = 0.0;
  // Integrator segment tables (max_active entries each), a ring of two so
  // the host can fill one while the device still reads the other; done[k]
  // marks the end of the last round that used table k.
  static constexpr int kRing = 2;
  khrr::IntSeg* seg_dev[kRing] = {};
  khrr::IntSeg* seg_pin[kRing] = {};
  hipEvent_t done[kRing] = {};
  bool recorded[kRing] = {};
  long long ring_pos = 0;
};

namespace {

// One batched gravity evaluation over `set`; flags may ask for the graph
// path (rounds repeat a composition until a system joins or leaves).
int gravity_over(khrr_server* s, const std::vector<Sim*>& set, float* acc, float* phi,
                 unsigned flags = 0u) {
  std::vector<int> begin, count;
  for (Sim* m : set) {
    begin.push_back(m->begin);
    count.push_back(m->n);
  }
  khrr_tune t = s->tune;
  t.i_per_lane = khrr::round_ipl(s->ctx, (int)set.size(), count.data(), s->tune.i_per_lane);
  return khrr::batched_segments(s->ctx, (int)set.size(), begin.data(), count.data(), s->cap,
                                s->pos32, s->h32, s->m32, acc, phi, flags, t);
}

int upload(khrr_server* s, Sim* m) {
  const size_t n = (size_t)m->n;
  khrr_ctx* ctx = s->ctx;
  KHRR_HIP(ctx, hipMemcpyAsync(m->d_pos(), m->pos.data(), 3 * n * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
  KHRR_HIP(ctx, hipMemcpyAsync(m->d_vel(), m->vel.data(), 3 * n * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
  KHRR_HIP(ctx, hipMemcpyAsync(m->d_mass(), m->mass.data(), n * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
  for (int c = 0; c < 3; ++c)
    KHRR_HIP(ctx, hipMemcpyAsync(s->pos32 + (size_t)c * s->cap + m->begin, m->pos32.data() + c * n,
                                 n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  KHRR_HIP(ctx, hipMemcpyAsync(s->m32 + m->begin, m->m32.data(), n * sizeof(float),
                               hipMemcpyHostToDevice, ctx->stream));
  KHRR_HIP(ctx, hipMemcpyAsync(s->h32 + m->begin, m->h32.data(), n * sizeof(float),
                               hipMemcpyHostToDevice, ctx->stream));
  // pageable sources: finish before the host copies are released
  KHRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int admit(khrr_server* s, std::vector<Sim*>& admitted) {
  while (!s->pending.empty() && (int)s->active.size() < s->max_active) {
    Sim* m = s->pending.front();
    const size_t foot = footprint(m->n);
    if (s->in_use + foot > s->budget) break;  // first come, first served: wait
    const int b = s->ranges.take(m->n);
    if (b < 0) break;
    if (hipMalloc(&m->dev, kStateDoublesPerBody * sizeof(double) * (size_t)m->n) != hipSuccess) {
      (void)hipGetLastError();
      m->dev = nullptr;
      s->ranges.give(b, m->n);
      break;
    }
    m->begin = b;
    if (const int rc = upload(s, m)) {
      // stays pending with nothing held, so a later round can admit it
      (void)hipFree(m->dev);
      m->dev = nullptr;
      m->begin = -1;
      s->ranges.give(b, m->n);
      return rc;
    }
    std::vector<double>().swap(m->pos);
    std::vector<double>().swap(m->vel);
    std::vector<double>().swap(m->mass);
    std::vector<float>().swap(m->pos32);
    std::vector<float>().swap(m->m32);
    std::vector<float>().swap(m->h32);
    m->state = State::kActive;
    s->in_use += foot;
    s->pending.pop_front();
    s->active.push_back(m);
    admitted.push_back(m);
  }
  if (admitted.empty()) return 0;
  // initial ACCEL + PHI (nbody.cu:1468-1481) and the t = 0 diagnostics
  KHRR_TRY(gravity_over(s, admitted, s->acc, s->phi));
  for (Sim* m : admitted) {
    KHRR_TRY(khrr_diag(s->ctx, m->n, m->d_pos(), m->d_vel(), m->d_mass(), s->phi + m->begin,
                       m->diag0));
    s->pairs += 2.0 * m->n * (double)m->n;
  }
  return 0;
}

int finish(khrr_server* s, const std::vector<Sim*>& done, long long* ids, int max_ids,
           int* n_ids) {
  if (done.empty()) return 0;
  khrr_ctx* ctx = s->ctx;
  KHRR_TRY(gravity_over(s, done, nullptr, s->phi));
  for (Sim* m : done) {
    KHRR_TRY(khrr_diag(ctx, m->n, m->d_pos(), m->d_vel(), m->d_mass(), s->phi + m->begin,
                       m->diag1));
    s->pairs += m->n * (double)m->n;
    const size_t G = (size_t)m->grid.size[0] * m->grid.size[1] * m->grid.size[2];
    if (G > s->bin_cells) {
      if (s->bin_count) KHRR_HIP(ctx, hipFree(s->bin_count));
      if (s->bin_mass) KHRR_HIP(ctx, hipFree(s->bin_mass));
      s->bin_count = nullptr;
      s->bin_mass = nullptr;
      s->bin_cells = 0;
      KHRR_HIP(ctx, hipMalloc(&s->bin_count, G * sizeof(int)));
      KHRR_HIP(ctx, hipMalloc(&s->bin_mass, G * sizeof(double)));
      s->bin_cells = G;
    }
    KHRR_TRY(khrr_p2mesh_bin(ctx, &m->grid, m->n, m->d_pos(), m->d_mass(), nullptr, nullptr,
                             s->bin_count, nullptr, nullptr, nullptr, s->bin_mass));
    m->cell_mass.resize(G);
    KHRR_HIP(ctx, hipMemcpyAsync(m->cell_mass.data(), s->bin_mass, G * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
    KHRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    KHRR_HIP(ctx, hipFree(m->dev));
    // released: out of the active set at once, so an error later in this
    // loop cannot leave a freed simulation active
    m->dev = nullptr;
    s->ranges.give(m->begin, m->n);
    s->in_use -= footprint(m->n);
    m->state = State::kFinished;
    s->active.erase(std::find(s->active.begin(), s->active.end(), m));
    if (*n_ids < max_ids && ids) ids[*n_ids] = m->id;
    ++*n_ids;
  }
  return 0;
}

}  // namespace

extern "C" {

int khrr_server_create(khrr_ctx* ctx, size_t hbm_budget, int max_active, khrr_server** out) {
  KHRR_TRY(khrr::enter(ctx));
  if (!out || max_active < 1) return khrr::fail(ctx, KHRR_EINVAL, "bad arguments");
  *out = nullptr;
  const size_t cap = hbm_budget / kBytesPerBody;
  if (cap < 1 || cap > (size_t)INT32_MAX / 4)
    return khrr::fail(ctx, KHRR_EINVAL, "budget %zu bytes gives %zu bodies", hbm_budget, cap);
  auto* s = new khrr_server();
  s->ctx = ctx;
  s->budget = hbm_budget;
  s->max_active = max_active;
  s->cap = (int)cap;
  if (hipMalloc(&s->arena, kArenaFloatsPerBody * sizeof(float) * cap) != hipSuccess) {
    (void)hipGetLastError();
    delete s;
    return khrr::fail(ctx, KHRR_ENOMEM, "server arena of %zu bodies", cap);
  }
  float* p = s->arena;
  s->pos32 = p;
  p += 3 * cap;
  s->m32 = p;
  p += cap;
  s->h32 = p;
  p += cap;
  s->acc = p;
  p += 3 * cap;
  s->acct = p;
  p += 3 * cap;
  s->phi = p;
  s->ranges.init((int)cap);
  bool ok = true;
  for (int k = 0; k < khrr_server::kRing && ok; ++k)
    ok = hipMalloc(&s->seg_dev[k], sizeof(khrr::IntSeg) * max_active) == hipSuccess &&
         hipHostMalloc(&s->seg_pin[k], sizeof(khrr::IntSeg) * max_active) == hipSuccess &&
         hipEventCreateWithFlags(&s->done[k], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    for (int k = 0; k < khrr_server::kRing; ++k) {
      if (s->seg_dev[k]) (void)hipFree(s->seg_dev[k]);
      if (s->seg_pin[k]) (void)hipHostFree(s->seg_pin[k]);
      if (s->done[k]) (void)hipEventDestroy(s->done[k]);
    }
    (void)hipFree(s->arena);
    delete s;
    return khrr::fail(ctx, KHRR_ENOMEM, "server tables for %d simulations", max_active);
  }
  if (ctx->has_batch_tuned) s->tune = ctx->batch_tuned;
  s->tune.j_split = 1;
  s->tune.symmetric = 0;
  *out = s;
  return 0;
}

void khrr_server_destroy(khrr_server* s) {
  if (!s) return;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  for (auto& kv : s->sims)
    if (kv.second->dev) (void)hipFree(kv.second->dev);
  if (s->arena) (void)hipFree(s->arena);
  for (int k = 0; k < khrr_server::kRing; ++k) {
    if (s->seg_dev[k]) (void)hipFree(s->seg_dev[k]);
    if (s->seg_pin[k]) (void)hipHostFree(s->seg_pin[k]);
    if (s->done[k]) (void)hipEventDestroy(s->done[k]);
  }
  if (s->bin_count) (void)hipFree(s->bin_count);
  if (s->bin_mass) (void)hipFree(s->bin_mass);
  delete s;
}

int khrr_server_set_round_pairs(khrr_server* s, double pairs) {
  if (!s) return khrr::fail(nullptr, KHRR_EINVAL, "null server");
  if (!(pairs >= 0.0)) return khrr::fail(s->ctx, KHRR_EINVAL, "round_pairs must be >= 0");
  s->round_pairs = pairs;
  return 0;
}

int khrr_server_submit(khrr_server* s, const khrr_request* req, long long* id) {
  if (!s) return khrr::fail(nullptr, KHRR_EINVAL, "null server");
  khrr_ctx* ctx = s->ctx;
  if (!req || !id) return khrr::fail(ctx, KHRR_EINVAL, "null argument");
  if (req->n < 1 || !req->pos || !req->vel || !req->mass || !req->eps)
    return khrr::fail(ctx, KHRR_EINVAL, "request needs n >= 1 and host arrays");
  if (req->steps < 0 || !(req->dt > 0.0)) return khrr::fail(ctx, KHRR_EINVAL, "bad dt/steps");
  // the result grid is binned when the simulation finishes: refuse now what
  // khrr_p2mesh_bin would refuse then
  KHRR_TRY(khrr::check_grid(ctx, &req->grid));
  if (req->n > s->cap || footprint(req->n) > s->budget)
    return khrr::fail(ctx, KHRR_ENOMEM, "request of %d bodies can never fit the budget", req->n);
  auto m = std::make_unique<Sim>();
  const size_t n = (size_t)req->n;
  m->id = s->next_id++;
  m->n = req->n;
  m->dt = req->dt;
  m->steps = req->steps;
  m->grid = req->grid;
  m->pos.assign(req->pos, req->pos + 3 * n);
  m->vel.assign(req->vel, req->vel + 3 * n);
  m->mass.assign(req->mass, req->mass + n);
  m->pos32.resize(3 * n);
  m->m32.resize(n);
  m->h32.resize(n);
  for (size_t i = 0; i < 3 * n; ++i) m->pos32[i] = (float)m->pos[i];
  for (size_t i = 0; i < n; ++i) {
    m->m32[i] = (float)m->mass[i];
    m->h32[i] = (float)(0.5 * (req->eps[i] * req->eps[i]));
  }
  *id = m->id;
  s->pending.push_back(m.get());
  s->sims.emplace(m->id, std::move(m));
  return 0;
}

int khrr_server_step(khrr_server* s, long long* done_ids, int max_done, int* n_done) {
  if (!s) return khrr::fail(nullptr, KHRR_EINVAL, "null server");
  khrr_ctx* ctx = s->ctx;
  KHRR_TRY(khrr::enter(ctx));
  if (!n_done) return khrr::fail(ctx, KHRR_EINVAL, "n_done is NULL");
  *n_done = 0;
  std::vector<Sim*> admitted;
  KHRR_TRY(admit(s, admitted));
  std::vector<Sim*> stepping, done;
  for (Sim* m : s->active) (m->done_steps < m->steps ? stepping : done).push_back(m);
  if (s->round_pairs > 0.0 && stepping.size() > 1) {
    auto left = [](const Sim* m) { return (double)m->n * m->n * (m->steps - m->done_steps); };
    std::stable_sort(stepping.begin(), stepping.end(), [&](const Sim* a, const Sim* b) {
      const double la = left(a), lb = left(b);
      return la != lb ? la < lb : a->id < b->id;
    });
    double pairs = 0.0;
    size_t k = 0;
    while (k < stepping.size() && (k == 0 || pairs < s->round_pairs)) {
      pairs += (double)stepping[k]->n * stepping[k]->n;
      ++k;
    }
    stepping.resize(k);
  }
  if (!stepping.empty()) {
    // predictor -> one batched gravity launch -> corrector (nbody.cu:1525-1582)
    const int slot = (int)(s->ring_pos++ % khrr_server::kRing);
    // the round that last used this table (two rounds ago) must be done
    // before its pinned copy is overwritten
    if (s->recorded[slot]) KHRR_HIP(ctx, hipEventSynchronize(s->done[slot]));
    khrr::IntSeg* pin = s->seg_pin[slot];
    khrr::IntSeg* seg = s->seg_dev[slot];
    int blocks = 0;
    for (size_t k = 0; k < stepping.size(); ++k) {
      Sim* m = stepping[k];
      pin[k] = khrr::IntSeg{m->dev, m->dt, m->n, m->begin, blocks, 0};
      blocks += (int)khrr::ceil_div(m->n, 256);
    }
    const int n_seg = (int)stepping.size();
    KHRR_HIP(ctx, hipMemcpyAsync(seg, pin, n_seg * sizeof(khrr::IntSeg), hipMemcpyHostToDevice,
                                 ctx->stream));
    KHRR_TRY(khrr::integrate_segments(ctx, n_seg, seg, blocks, 0, s->acc, nullptr, s->pos32,
                                      s->cap));
    KHRR_TRY(gravity_over(s, stepping, s->acct, nullptr, KHRR_GRAPH));
    KHRR_TRY(khrr::integrate_segments(ctx, n_seg, seg, blocks, 1, s->acct, s->acc, nullptr,
                                      s->cap));
    KHRR_HIP(ctx, hipEventRecord(s->done[slot], ctx->stream));
    s->recorded[slot] = true;
    for (Sim* m : stepping) {
      ++m->done_steps;
      s->pairs += m->n * (double)m->n;
      if (m->done_steps >= m->steps) done.push_back(m);
    }
    ++s->rounds;
  }
  KHRR_TRY(finish(s, done, done_ids, max_done, n_done));
  // rounds that only step return with their launches queued; admission and
  // finishing have already waited for the device
  if (!admitted.empty() || !done.empty()) KHRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int khrr_server_result(khrr_server* s, long long id, double* diag0, double* diag1,
                       double* cell_mass, size_t cell_mass_len) {
  if (!s) return khrr::fail(nullptr, KHRR_EINVAL, "null server");
  auto it = s->sims.find(id);
  if (it == s->sims.end()) return khrr::fail(s->ctx, KHRR_EINVAL, "unknown request %lld", id);
  Sim* m = it->second.get();
  if (m->state != State::kFinished)
    return khrr::fail(s->ctx, KHRR_ESTATE, "request %lld has not finished", id);
  if (cell_mass && cell_mass_len != m->cell_mass.size())
    return khrr::fail(s->ctx, KHRR_EINVAL, "cell_mass holds %zu cells, grid has %zu",
                      cell_mass_len, m->cell_mass.size());
  if (diag0) std::memcpy(diag0, m->diag0, sizeof(m->diag0));
  if (diag1) std::memcpy(diag1, m->diag1, sizeof(m->diag1));
  if (cell_mass) std::memcpy(cell_mass, m->cell_mass.data(), cell_mass_len * sizeof(double));
  s->sims.erase(it);
  return 0;
}

int khrr_server_stats(khrr_server* s, khrr_server_info* out) {
  if (!s || !out) return khrr::fail(nullptr, KHRR_EINVAL, "null argument");
  int finished = 0;
  for (auto& kv : s->sims) finished += kv.second->state == State::kFinished;
  out->active = (int)s->active.size();
  out->pending = (int)s->pending.size();
  out->finished = finished;
  out->rounds = s->rounds;
  out->pair_interactions = s->pairs;
  out->bytes_in_use = s->in_use;
  out->bytes_budget = s->budget;
  return 0;
}

}  // extern "C"
