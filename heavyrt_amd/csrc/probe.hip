// probe.hip -- hrt_issue_probe: a calibration kernel for issue-rate
// measurements.  Eight waves on every SIMD run long chains of independent
// FP32 min / max / multiply / compare-and-select instructions (the
// operations a slab test issues) with no memory traffic in the loop, so the
// VALU instructions it issues per cycle (rocprofv3 SQ_INSTS_VALU against
// GRBM_GUI_ACTIVE) are the most a SIMD issues for such code.  A traversal
// kernel's count per cycle divided by the probe's is its issue fraction.
#include <hip/hip_runtime.h>

#include "hrt_internal.h"

namespace {

__global__ __launch_bounds__(64) void k_issue_probe(float* __restrict__ sink, int iters, float b, float c) {
  float a0 = float(threadIdx.x), a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
  for (int i = 0; i < iters; i++) {
    a0 = fminf(a0 * b, c);
    a1 = fmaxf(a1 * b, c);
    a2 = (a2 < c) ? a2 * b : c;
    a3 = fminf(fmaxf(a3, b), c);
  }
  if (threadIdx.x == 0) sink[blockIdx.x] = a0 + a1 + a2 + a3;  // keeps the chains alive
}

}  // namespace

extern "C" hrt_status hrt_issue_probe(hrt_ctx* ctx, void* stream) {
  if (!ctx) return hrt::fail(HRT_ERR_INVALID_ARG, "context is null");
  hipError_t e = hipSetDevice(ctx->device);
  int cus = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
  if (e != hipSuccess) return hrt::fail_hip(e, "issue probe");
  hipStream_t hs = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  // 4 SIMDs per CU, 8 waves per SIMD, one wave per workgroup
  const uint32_t blocks = uint32_t(cus) * 32u;
  float* sink = nullptr;
  if ((e = hipMallocAsync(reinterpret_cast<void**>(&sink), sizeof(float) * blocks, hs)) != hipSuccess)
    return hrt::fail_hip(e, "issue probe");
  k_issue_probe<<<blocks, 64, 0, hs>>>(sink, 100000, 0.999f, 0.001f);
  e = hipGetLastError();
  const hipError_t f = hipFreeAsync(sink, hs);
  if (e == hipSuccess) e = f;
  if (e != hipSuccess) return hrt::fail_hip(e, "issue probe");
  return HRT_OK;
}
