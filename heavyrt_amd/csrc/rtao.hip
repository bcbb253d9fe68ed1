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

// The direction bin of AO sample k, from the same two PCG32 draws as ao_ray
// but with the hardware's approximate sine, cosine, square root and
// reciprocal: a bin only steers which rays share a wave, so a sample near a
// bin edge may land in either neighbour without changing any answer, and the
// approximations take a fraction of ao_ray's instructions.
__device__ inline uint32_t ao_bin_fast(const AoFrame& f, uint64_t k, uint64_t seed) {
  if (f.org.w == 0.0f) return ao_dir_bin(make_float3(0.f, 1.f, 0.f));
  const uint64_t inc = (k << 1u) | 1u;
  uint64_t st = 0;
  (void)ao_pcg32_step(st, inc);
  st += seed;
  (void)ao_pcg32_step(st, inc);
  const float r1 = float(ao_pcg32_step(st, inc) >> 8) * (1.0f / 16777216.0f);
  const float r2 = float(ao_pcg32_step(st, inc) >> 8) * (1.0f / 16777216.0f);
  const float sphi = __builtin_amdgcn_sinf(r1), cphi = __builtin_amdgcn_cosf(r1);  // of 2 pi r1
  const float cth = __builtin_amdgcn_sqrtf(1.0f - r2), sth = __builtin_amdgcn_sqrtf(r2);
  const float3 n = make_float3(f.n.x, f.n.y, f.n.z);
  const float sgn = copysignf(1.0f, n.z);
  const float a = f.n.w;
  const float b = n.x * n.y * a;
  const float lx = sth * cphi, ly = sth * sphi;
  return ao_dir_bin(make_float3((1.0f + sgn * n.x * n.x * a) * lx + b * ly + n.x * cth,
                                sgn * b * lx + (sgn + n.y * n.y * a) * ly + n.y * cth,
                                -sgn * n.x * lx - n.y * ly + n.z * cth));
}

// Direction order of the AO samples, one wave per run of kAoOrderRun
// consecutive samples: order[run_base + r] = the offset within the run of the
// sample traced r-th, the run's samples sorted stably by ao_dir_bin.  Lane l
// holds samples l + 64 j (j = 0..15); slot by slot, the lanes that share a
// bin find each other with five ballots (one per bin bit), so every sample's
// place is its bin's start + the bin's samples in earlier slots + the lanes
// before it in this slot.
__global__ __launch_bounds__(64) void k_ao_order(const AoFrame* __restrict__ frames, uint64_t n, uint32_t spp,
                                                 uint64_t seed, uint16_t* __restrict__ order) {
  constexpr uint32_t kSlots = kAoOrderRun / 64;
  __shared__ uint32_t count[32], cursor[32];
  const uint32_t lane = threadIdx.x;
  const uint64_t base = uint64_t(blockIdx.x) * kAoOrderRun;
  if (lane < 32) count[lane] = 0u;
  __syncthreads();
  uint32_t bin[kSlots];
#pragma unroll
  for (uint32_t j = 0; j < kSlots; j++) {
    const uint64_t i = base + lane + 64u * j;
    bin[j] = kAoBins;  // past the end: never placed
    if (i < n) {
      bin[j] = ao_bin_fast(frames[uint32_t(i) / spp], i, seed);
      atomicAdd(&count[bin[j]], 1u);
    }
  }
  __syncthreads();
  if (lane == 0) {
    uint32_t s = 0;
    for (uint32_t b = 0; b < kAoBins; b++) {
      cursor[b] = s;
      s += count[b];
    }
  }
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < kSlots; j++) {
    const uint32_t b = bin[j];
    uint64_t peers = __ballot(b < kAoBins);
#pragma unroll
    for (uint32_t k = 0; k < 5; k++) {
      const uint64_t m = __ballot((b >> k) & 1u);
      peers &= ((b >> k) & 1u) ? m : ~m;
    }
    if (b < kAoBins) {
      const uint32_t at = cursor[b] + lanes_below(peers);
      order[base + at] = uint16_t(lane + 64u * j);
    }
    __syncthreads();
    // the lowest lane of each group moves its bin's cursor past the group
    if (b < kAoBins && lanes_below(peers) == 0) cursor[b] += uint32_t(__popcll(peers));
    __syncthreads();
  }
}

}  // namespace

hipError_t rtao_order(const AoFrame* frames, uint64_t npix, uint32_t spp, uint64_t seed,
                      uint16_t* order, hipStream_t stream) {
  const uint64_t n = npix * spp;
  if (n == 0) return hipSuccess;
  const uint32_t runs = uint32_t((n + kAoOrderRun - 1) / kAoOrderRun);
  k_ao_order<<<runs, 64, 0, stream>>>(frames, n, spp, seed, order);
  return hipGetLastError();
}

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
