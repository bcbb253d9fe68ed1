// rtao.hip -- device-side ambient-occlusion ray generation: one lane makes
// one AO ray in HBM (k_ao_rays), or one pixel's frame for the traversal that
// makes its rays in registers (k_ao_frames).  The arithmetic is in ao.h.
#include <hip/hip_runtime.h>

#include "ao.h"
#include "hrt_internal.h"

namespace hrt {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_ao_rays(const Ray* __restrict__ prim_rays,
                                                    const uint2* __restrict__ prim_hits,
                                                    const float4* __restrict__ verts,
                                                    const uint32_t* __restrict__ idx, uint32_t ntris, uint64_t npix,
                                                    uint32_t spp, float tmax, uint64_t seed, Ray* __restrict__ out) {
  const uint64_t k = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (k >= npix * spp) return;
  const AoFrame f = ao_frame(prim_rays, prim_hits, verts, idx, ntris, k / spp);
  out[k] = ao_ray(f, k, tmax, seed);
}

__global__ __launch_bounds__(kBlock) void k_ao_frames(const Ray* __restrict__ prim_rays,
                                                      const uint2* __restrict__ prim_hits,
                                                      const float4* __restrict__ verts,
                                                      const uint32_t* __restrict__ idx, uint32_t ntris,
                                                      uint64_t npix, AoFrame* __restrict__ out) {
  const uint64_t p = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= npix) return;
  out[p] = ao_frame(prim_rays, prim_hits, verts, idx, ntris, p);
}

}  // namespace

hipError_t rtao_make_rays(const Blas& b, const Ray* prim_rays, const void* prim_hits_compact, uint64_t npix,
                          uint32_t spp, float tmax, uint64_t seed, Ray* out, hipStream_t stream) {
  const uint64_t n = npix * spp;
  if (n == 0) return hipSuccess;
  const uint32_t blocks = uint32_t((n + kBlock - 1) / kBlock);
  k_ao_rays<<<blocks, kBlock, 0, stream>>>(prim_rays, static_cast<const uint2*>(prim_hits_compact), b.verts, b.idx,
                                           b.ntris, npix, spp, tmax, seed, out);
  return hipGetLastError();
}

hipError_t rtao_frames(const Blas& b, const Ray* prim_rays, const void* prim_hits_compact, uint64_t npix,
                       AoFrame* out, hipStream_t stream) {
  if (npix == 0) return hipSuccess;
  const uint32_t blocks = uint32_t((npix + kBlock - 1) / kBlock);
  k_ao_frames<<<blocks, kBlock, 0, stream>>>(prim_rays, static_cast<const uint2*>(prim_hits_compact), b.verts,
                                             b.idx, b.ntris, npix, out);
  return hipGetLastError();
}

}  // namespace hrt
