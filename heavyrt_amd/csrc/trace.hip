// trace.hip -- BVH2 nearest-hit / any-hit traversal for gfx950.
//
// Semantics follow BVH2CommonRT::BVH2TraverseC32 (core/BVH2CommonRT.cpp:49-141)
// step for step: start at the root's children, test both boxes of a pair,
// descend into the nearer hit child (the left one on equal entry distance,
// :106) and defer the other, intersect every triangle of a leaf with the
// reference's Moller-Trumbore (first strictly closer hit wins, :37), pop when
// a leaf is done or both children miss.  Box culling uses the current hit
// distance (:97-98).  A ray whose tFar is <= 0 is an any-hit query with
// tFar = -w that stops at its first hit (BVH4HalfRT.cpp:260-262, 82).
//
// Layout: NodePair (64 B, both children of one internal node) and TriPre
// (48 B, A / B-A / C-A / original id), both in HBM; the per-lane stack of
// deferred children lives in LDS, 32 entries, interleaved by lane.
#include <hip/hip_runtime.h>

#include "hrt_internal.h"

namespace hrt {
namespace {

constexpr int kTraceBlock = 256;

struct HitOutFull {  // CRT_Hit (CrossRT.h:20-27)
  float t;
  uint32_t prim, inst, geom;
  float coords[4];
};

template <bool kForceAny, bool kStats>
__global__ __launch_bounds__(kTraceBlock) void k_trace_bvh2(const NodePair* __restrict__ pairs,
                                                            const TriPre* __restrict__ tris,
                                                            const Ray* __restrict__ rays, uint64_t n, void* out,
                                                            int format, uint32_t* __restrict__ stats) {
  __shared__ uint32_t stack_mem[kStackDepth * kTraceBlock];
  const uint64_t i = uint64_t(blockIdx.x) * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t* stack = stack_mem + threadIdx.x;  // entry k at stack[k * kTraceBlock]

  const Ray r = rays[i];
  const float3 o = make_float3(r.pos_near.x, r.pos_near.y, r.pos_near.z);
  const float3 d = make_float3(r.dir_far.x, r.dir_far.y, r.dir_far.z);
  const float tnear = r.pos_near.w;
  const bool any = kForceAny || (r.dir_far.w <= 0.0f);
  const float tfar = kForceAny ? fabsf(r.dir_far.w) : (any ? -r.dir_far.w : r.dir_far.w);
  const float3 inv = safe_inverse(d);

  HitRec h;
  h.t = tfar;
  h.prim = kInvalid;
  h.u = 0.0f;
  h.v = 0.0f;
  uint32_t npairs = 0, ntests = 0;

  int top = 0;
  uint32_t cur = 0;  // pair index of the root's children
  while (top >= 0) {
    bool pop;
    if ((cur & kLeafBit) == 0) {
      const float4* pp = reinterpret_cast<const float4*>(pairs + cur);
      const float4 lmin = pp[0], lmax = pp[1], rmin = pp[2], rmax = pp[3];
      if (kStats) npairs++;
      const float2 t0 = slab(o, inv, lmin.x, lmin.y, lmin.z, lmax.x, lmax.y, lmax.z);
      const float2 t1 = slab(o, inv, rmin.x, rmin.y, rmin.z, rmax.x, rmax.y, rmax.z);
      const bool h0 = (t0.x <= t0.y) && (t0.y >= tnear) && (t0.x <= h.t);
      const bool h1 = (t1.x <= t1.y) && (t1.y >= tnear) && (t1.x <= h.t);
      const uint32_t c0 = __float_as_uint(lmin.w), c1 = __float_as_uint(rmin.w);
      pop = !h0 && !h1;
      cur = h0 ? c0 : c1;
      if (h0 && h1) {
        const bool left_first = t0.x <= t1.x;
        cur = left_first ? c0 : c1;
        stack[top * kTraceBlock] = left_first ? c1 : c0;
        top++;
      }
    } else {
      pop = true;
      if (cur != kInvalid) {
        const uint32_t s = leaf_start(cur), c = leaf_count(cur);
        if (kStats) ntests += c;
        for (uint32_t k = s; k < s + c; k++) mt_test(o, d, tnear, tris[k], h);
        if (any && h.prim != kInvalid) break;
      }
    }
    if (pop) {
      top--;
      cur = stack[max(top, 0) * kTraceBlock];
    }
  }

  if (kStats) {
    stats[2 * i + 0] = npairs;
    stats[2 * i + 1] = ntests;
  }
  if (kForceAny) {
    static_cast<uint8_t*>(out)[i] = (h.prim != kInvalid) ? 1 : 0;
  } else if (format == kHitCompact) {
    static_cast<uint2*>(out)[i] = make_uint2(__float_as_uint(h.t), h.prim);
  } else {
    const bool hit = h.prim != kInvalid;
    float4* o4 = reinterpret_cast<float4*>(static_cast<HitOutFull*>(out) + i);
    o4[0] = make_float4(h.t, __uint_as_float(h.prim), __uint_as_float(hit ? 0u : kInvalid),
                        __uint_as_float(hit ? 0u : kInvalid));
    o4[1] = make_float4(h.u, h.v, 0.0f, 0.0f);
  }
}

}  // namespace

hipError_t trace_blas(const Blas& b, const TraceArgs& a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  const uint32_t blocks = uint32_t((a.n + kTraceBlock - 1) / kTraceBlock);
  if (a.anyhit) {
    if (a.stats)
      k_trace_bvh2<true, true><<<blocks, kTraceBlock, 0, stream>>>(b.pairs, b.tris, a.rays, a.n, a.out, a.format,
                                                                    a.stats);
    else
      k_trace_bvh2<true, false><<<blocks, kTraceBlock, 0, stream>>>(b.pairs, b.tris, a.rays, a.n, a.out, a.format,
                                                                     nullptr);
  } else {
    if (a.stats)
      k_trace_bvh2<false, true><<<blocks, kTraceBlock, 0, stream>>>(b.pairs, b.tris, a.rays, a.n, a.out, a.format,
                                                                     a.stats);
    else
      k_trace_bvh2<false, false><<<blocks, kTraceBlock, 0, stream>>>(b.pairs, b.tris, a.rays, a.n, a.out, a.format,
                                                                      nullptr);
  }
  return hipGetLastError();
}

}  // namespace hrt
