// noalloc.hip -- hrt_trace_any_noalloc: hrt_trace_any without any allocation.
//
// The persistent any-hit kernels need scratch beyond their inputs and
// outputs: a ray-queue counter and a per-lane stack spill area.
// hrt_trace_any takes it from the stream-ordered pool on every call; this
// entry point takes it from `scratch`, a file-scope device array of
// kScratchWords words (96 MiB + 256 B) that each GPU the library runs on holds
// until the process ends.  Its device address is looked up once per device
// by noalloc_prepare (hrt_init, with the device current) and kept in
// g_scratch; host code never passes `scratch` itself to a kernel.  The queue
// counter is cleared by a fill kernel before every launch.  Every call on one
// GPU shares that GPU's array: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hrt_internal.h"

namespace {

// queue counter (64 words) + stacks for 8,192 workgroups of 64 lanes x 48
// spill words (the BVH4 kernels need kStackDepth4 - kShortStack = 47, the
// BVH2 kernels 16)
constexpr size_t kScratchWords = 64 + size_t(8192) * 64 * 48;
static_assert(kScratchWords >= 64 + size_t(8192) * 64 * (hrt::kStackDepth4 - hrt::kShortStack),
              "the scratch array holds stacks for 8,192 workgroups");

__device__ __align__(256) uint32_t scratch[kScratchWords];

struct DeviceScratch {
  std::once_flag once;
  uint32_t* ptr = nullptr;
  hipError_t err = hipSuccess;
};
constexpr int kMaxDevices = 64;
DeviceScratch g_scratch[kMaxDevices];

}  // namespace

namespace hrt {

hipError_t noalloc_prepare(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  DeviceScratch& d = g_scratch[device];
  std::call_once(d.once, [&] {
    d.err = hipGetSymbolAddress(reinterpret_cast<void**>(&d.ptr), HIP_SYMBOL(scratch));
    if (d.err == hipSuccess && !d.ptr) d.err = hipErrorInvalidSymbol;
    if (d.err == hipSuccess) warm_any_hit_launchers();
  });
  return d.err;
}

}  // namespace hrt

extern "C" {

hrt_status hrt_trace_any_noalloc(hrt_scene* s, const float* rays, size_t n, uint8_t* out, void* stream) {
  if (!s) return hrt::fail(HRT_ERR_INVALID_ARG, "scene is null");
  const int device = hrt::scene_device(s);
  if (device < 0 || device >= kMaxDevices || !g_scratch[device].ptr)
    return hrt::fail(HRT_ERR_INVALID_ARG, "no scratch on this device (create a context on it with hrt_init first)");
  return hrt::trace_any_in_workspace(s, rays, n, out, stream, g_scratch[device].ptr, kScratchWords);
}

}  // extern "C"
