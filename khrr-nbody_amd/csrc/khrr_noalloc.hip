// khrr_noalloc.hip -- khrr_p2mesh_bin_noalloc, the binning entry point that
// allocates nothing.
//
// khrr_p2mesh_bin takes its scratch (cell indices, the radix sort's key and
// value buffers, tile histograms, block bounds) from the context's arena,
// which grows with hipMalloc on demand.  Callers that cannot allocate --
// code under HIP graph capture, tight serving loops -- use this variant: the
// same kernels, with the scratch in one fixed device array of this file,
// sized for up to kMaxBodies bodies on up to kMaxCells cells.
//
// The array's name is a host-side stand-in, not a device address.  Each
// device gets its own copy of it at its own address, looked up once per
// device with hipGetSymbolAddress when a context is created there
// (noalloc_init), and only that address is handed to kernels.
//
// Every call on one device shares that device's array: calls that overlap
// on one device (from several contexts, streams or host threads) race and
// must be serialised by the caller.  No word of the array needs to start
// at zero: every call writes each word it reads before reading it.
#include <mutex>

#include "khrr_internal.h"

namespace {

constexpr long long kMaxBodies = 1ll << 20;  // 1048576 bodies
constexpr long long kMaxCells = 1ll << 24;   // 256^3 cells
constexpr size_t kScratchBytes = khrr::bin_scratch_need(kMaxBodies, kMaxCells);
static_assert(kScratchBytes < (size_t(32) << 20), "fixed binning scratch stays below 32 MiB");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace khrr {

int noalloc_init(khrr_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return fail(ctx, KHRR_EINVAL, "device %d beyond the %d the fixed scratch table covers",
                ctx->device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[ctx->device]) return 0;
  KHRR_HIP(ctx, hipSetDevice(ctx->device));
  void* p = nullptr;
  // ::scratch, the array: inside namespace khrr the bare name is the
  // context-arena function khrr::scratch
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
  if (e != hipSuccess || !p)
    return fail(ctx, KHRR_EDEVICE, "hipGetSymbolAddress of the fixed binning scratch failed: %s",
                hipGetErrorString(e));
  g_scratch[ctx->device] = p;
  return 0;
}

}  // namespace khrr

extern "C" int khrr_p2mesh_bin_noalloc(khrr_ctx* ctx, const khrr_grid* grid, int n,
                                       const double* pos, const double* mass, int* cell_id,
                                       int* id_in_cell, int* count, int* offset,
                                       int* start_id_ref, int* block_incl, double* cell_mass) {
  using khrr::fail;
  KHRR_TRY(khrr::enter(ctx));
  KHRR_TRY(khrr::bin_validate(ctx, grid, n, pos, mass, count, cell_mass));
  const long long G = (long long)grid->size[0] * grid->size[1] * grid->size[2];
  if (n > kMaxBodies || G > kMaxCells || khrr::bin_scratch_need(n, G) > kScratchBytes)
    return fail(ctx, KHRR_EINVAL,
                "khrr_p2mesh_bin_noalloc: %d bodies on %lld cells exceed its fixed scratch "
                "(at most %lld bodies and %lld cells); use khrr_p2mesh_bin",
                n, G, kMaxBodies, kMaxCells);
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    base = g_scratch[ctx->device];
  }
  if (!base)
    return fail(ctx, KHRR_ESTATE, "fixed binning scratch not set up on device %d", ctx->device);
  // no timers: their events come from a pool that may create one
  return khrr::bin_launch(ctx, grid, n, pos, mass, cell_id, id_in_cell, count, offset,
                          start_id_ref, block_incl, cell_mass, base, false);
}
