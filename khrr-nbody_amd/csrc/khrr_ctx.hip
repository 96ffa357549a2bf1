// khrr_ctx.hip -- context lifecycle, error reporting, scratch arena and
// kernel timers.  Replaces the RR::CUDA plumbing the reference links
// (CuDarray / CuCall / CuTimer / CuDeviceSync, used at nbody.cu:141-200,
// 1428-1441, 1447-1520), which carries no arithmetic.
#include "khrr_internal.h"

#include <cstring>

namespace khrr {

std::string& thread_error() {
  static thread_local std::string msg;
  return msg;
}

int fail(khrr_ctx* ctx, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx)
    ctx->err = buf;
  else
    thread_error() = buf;
  return code;
}

int enter(khrr_ctx* ctx) {
  if (!ctx) return fail(nullptr, KHRR_EINVAL, "null context");
  KHRR_HIP(ctx, hipSetDevice(ctx->device));
  return 0;
}

int scratch(khrr_ctx* ctx, size_t bytes, void** base) {
  if (bytes > ctx->scratch_bytes) {
    if (ctx->scratch) {
      // The old arena may still be in use by queued work.
      KHRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
      KHRR_HIP(ctx, hipFree(ctx->scratch));
      ctx->scratch = nullptr;
      ctx->scratch_bytes = 0;
    }
    size_t want = bytes + bytes / 4 + (1u << 20);
    if (hipMalloc(&ctx->scratch, want) != hipSuccess) {
      (void)hipGetLastError();  // clear it so a later launch check does not report it
      ctx->scratch = nullptr;
      return fail(ctx, KHRR_ENOMEM, "scratch allocation of %zu bytes failed", want);
    }
    ctx->scratch_bytes = want;
  }
  *base = ctx->scratch;
  return 0;
}

int aux_stream(khrr_ctx* ctx) {
  if (ctx->aux_stream) return 0;
  KHRR_HIP(ctx, hipEventCreateWithFlags(&ctx->fork_event, hipEventDisableTiming));
  KHRR_HIP(ctx, hipEventCreateWithFlags(&ctx->join_event, hipEventDisableTiming));
  KHRR_HIP(ctx, hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
  return 0;
}

int check_launch(khrr_ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(ctx, KHRR_EDEVICE, "launch of %s failed: %s", what, hipGetErrorString(e));
  return 0;
}

static hipEvent_t pool_get(khrr_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

Timed::Timed(khrr_ctx* ctx, int cls) : ctx_(ctx), cls_(cls) {
  if (ctx_ && ctx_->timing) {
    start_ = pool_get(ctx_);
    if (start_ && hipEventRecord(start_, ctx_->stream) != hipSuccess) {
      ctx_->event_pool.push_back(start_);
      start_ = nullptr;
    }
  }
}

Timed::~Timed() {
  if (!start_) return;
  hipEvent_t stop = pool_get(ctx_);
  if (!stop) return;
  if (hipEventRecord(stop, ctx_->stream) != hipSuccess) {
    ctx_->event_pool.push_back(start_);
    ctx_->event_pool.push_back(stop);
    return;
  }
  ctx_->pending.push_back(TimerPair{start_, stop, cls_});
}

static int harvest(khrr_ctx* ctx) {
  for (auto& p : ctx->pending) {
    KHRR_HIP(ctx, hipEventSynchronize(p.stop));
    float ms = 0.f;
    KHRR_HIP(ctx, hipEventElapsedTime(&ms, p.start, p.stop));
    ctx->launches[p.cls] += 1;
    ctx->total_ms[p.cls] += ms;
    ctx->event_pool.push_back(p.start);
    ctx->event_pool.push_back(p.stop);
  }
  ctx->pending.clear();
  return 0;
}

}  // namespace khrr

extern "C" {

int khrr_abi_version(void) { return KHRR_ABI_VERSION; }

int khrr_ctx_create(int device, khrr_ctx** out) {
  if (!out) return khrr::fail(nullptr, KHRR_EINVAL, "out is null");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0)
    return khrr::fail(nullptr, KHRR_EDEVICE, "no HIP device available (%s)",
                      hipGetErrorString(e));
  if (device < 0 || device >= ndev)
    return khrr::fail(nullptr, KHRR_EINVAL, "device %d out of range [0,%d)", device, ndev);
  if (hipSetDevice(device) != hipSuccess)
    return khrr::fail(nullptr, KHRR_EDEVICE, "hipSetDevice(%d) failed", device);
  khrr_ctx* ctx = new khrr_ctx();
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    if (prop.multiProcessorCount > 0) ctx->cu_count = prop.multiProcessorCount;
    ctx->arch = prop.gcnArchName;
  }
  if (hipMalloc(&ctx->dev_words, 256) != hipSuccess) {
    delete ctx;
    return khrr::fail(nullptr, KHRR_ENOMEM, "context word allocation failed");
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&ctx->host_stage), 4096) != hipSuccess) {
    (void)hipFree(ctx->dev_words);
    delete ctx;
    return khrr::fail(nullptr, KHRR_ENOMEM, "pinned staging allocation failed");
  }
  // The second stream exists from the start, so that no call creates it
  // while its caller captures the context's stream into a graph.
  if (khrr::noalloc_init(ctx) != 0 || khrr::aux_stream(ctx) != 0) {
    const std::string msg = ctx->err;
    if (ctx->aux_stream) (void)hipStreamDestroy(ctx->aux_stream);
    if (ctx->fork_event) (void)hipEventDestroy(ctx->fork_event);
    if (ctx->join_event) (void)hipEventDestroy(ctx->join_event);
    (void)hipHostFree(ctx->host_stage);
    (void)hipFree(ctx->dev_words);
    delete ctx;
    return khrr::fail(nullptr, KHRR_EDEVICE, "%s", msg.c_str());
  }
  *out = ctx;
  return 0;
}

void khrr_ctx_destroy(khrr_ctx* ctx) {
  if (!ctx) return;
  // Teardown is best effort: errors here have no caller to report to.
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& p : ctx->pending) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  for (auto& g : ctx->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    if (g.dev) (void)hipFree(g.dev);
  }
  if (ctx->capture_stream) (void)hipStreamDestroy(ctx->capture_stream);
  if (ctx->aux_stream) {
    (void)hipStreamSynchronize(ctx->aux_stream);
    (void)hipStreamDestroy(ctx->aux_stream);
  }
  if (ctx->fork_event) (void)hipEventDestroy(ctx->fork_event);
  if (ctx->join_event) (void)hipEventDestroy(ctx->join_event);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->batch_dev) (void)hipFree(ctx->batch_dev);
  if (ctx->batch_pinned) (void)hipHostFree(ctx->batch_pinned);
  if (ctx->dev_words) (void)hipFree(ctx->dev_words);
  if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
  delete ctx;
}

int khrr_ctx_set_stream(khrr_ctx* ctx, void* stream) {
  if (!ctx) return khrr::fail(nullptr, KHRR_EINVAL, "null context");
  ctx->stream = static_cast<hipStream_t>(stream);
  return 0;
}

const char* khrr_last_error(const khrr_ctx* ctx) {
  return ctx ? ctx->err.c_str() : khrr::thread_error().c_str();
}

int khrr_ctx_sync(khrr_ctx* ctx) {
  KHRR_TRY(khrr::enter(ctx));
  KHRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int khrr_ctx_set_timing(khrr_ctx* ctx, int enable) {
  if (!ctx) return khrr::fail(nullptr, KHRR_EINVAL, "null context");
  ctx->timing = enable != 0;
  return 0;
}

int khrr_ctx_kernel_time(khrr_ctx* ctx, int cls, long* launches, double* total_ms) {
  KHRR_TRY(khrr::enter(ctx));
  if (cls < 0 || cls >= KHRR_K_COUNT) return khrr::fail(ctx, KHRR_EINVAL, "bad kernel class %d", cls);
  KHRR_TRY(khrr::harvest(ctx));
  if (launches) *launches = ctx->launches[cls];
  if (total_ms) *total_ms = ctx->total_ms[cls];
  return 0;
}

int khrr_ctx_reset_timing(khrr_ctx* ctx) {
  KHRR_TRY(khrr::enter(ctx));
  KHRR_TRY(khrr::harvest(ctx));
  for (int i = 0; i < KHRR_K_COUNT; ++i) {
    ctx->launches[i] = 0;
    ctx->total_ms[i] = 0;
  }
  return 0;
}

}  // extern "C"
