// Allocation-free form of the one-shot kirchhoff_gpu solve, for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  Every device buffer of the solve
// (staged inputs, resident-form rays and slots, FP32 records, tiles, partial
// sums, launch metadata, the merged output) is carved from one file-scope
// device array of kScratchBytes; nothing is allocated or freed.
//
// Each GPU holds its own copy of the array for the life of the process.  Its
// address on a device is looked up once, with that device current, when a
// context is created on it (noalloc_scratch), and host code only ever uses
// that address.  Every call on one GPU shares the array, which keeps the
// last call's contents: the solve writes every byte it later reads (the
// queue counters are cleared by a fill kernel, the instance scratch by the
// centroid kernel, partial sums and the output in full), so nothing has to
// start at zero.  Calls through one context are serialised by its lock;
// calls through different contexts on one device must not overlap.
#include <hip/hip_runtime.h>

#include <mutex>

#include "runtime.h"

namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
}  // namespace

// The scratch block.  In host code this name is a host-side stand-in, not
// device memory: host code uses only the per-device address from
// hipGetSymbolAddress below.
static __device__ __align__(256) char scratch[kScratchBytes];

namespace xrtk {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_looked_up[kMaxDevices] = {};
}  // namespace

void* noalloc_scratch(int device, size_t* bytes) {
  *bytes = 0;
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_looked_up[device]) {
    // the caller has made `device` current: this is its copy of the array
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
    }
    g_scratch_ptr[device] = p;
    g_scratch_looked_up[device] = true;
  }
  if (g_scratch_ptr[device]) *bytes = kScratchBytes;
  return g_scratch_ptr[device];
}

}  // namespace xrtk

extern "C" int xrtk_kirchhoff_gpu_f32_noalloc(xrtk_ctx* ctx, const XrtkGpuRay* rays,
                                              size_t n_rays, const XrtkGpuPixel* pixels,
                                              size_t n_pixels, double* out4,
                                              const XrtkOpts* opts) {
  if (!ctx) return XRTK_ENODEV;
  if (!ctx->scratch) return XRTK_ELAUNCH;  // the address lookup failed at set-up
  xrtk::Arena arena;
  arena.base = static_cast<char*>(ctx->scratch);
  arena.cap = ctx->scratch_bytes;
  xrtk::Workspace ws;
  ws.use_arena(&arena);
  // a solve that needs more than the block fails at its first buffer that
  // does not fit, with XRTK_ENOMEM
  return xrtk::kirchhoff_f32_ws(ctx, ws, rays, n_rays, pixels, n_pixels, out4, opts);
}
