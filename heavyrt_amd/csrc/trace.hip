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
//
// Execution: a persistent wave64 kernel.  Each wave keeps 64 rays in flight;
// when enough lanes have finished, the wave counts its idle lanes with one
// ballot, takes that many consecutive ray indices from a global queue with a
// single atomic and refills exactly those lanes (ballot-based compaction), so
// long rays never hold a wave of finished lanes.
//
// Nearest-hit keeps the reference's exact order: a lane that reaches a leaf
// intersects it before it tests any further node.  Any-hit only has to decide
// whether some triangle is hit, which does not depend on the order, so it
// traverses speculatively (Aila & Laine, "Understanding the Efficiency of Ray
// Traversal on GPUs", HPG 2009): a lane that has found a leaf parks it and
// keeps descending until all but a few lanes of the wave hold a leaf, then
// those lanes intersect their leaves together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "hrt_internal.h"

namespace hrt {
namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 1;
constexpr int kTraceBlock = kWave * kWavesPerBlock;
constexpr int kRefillAt = 16;  // refill when at least this many lanes are idle
constexpr uint32_t kClaim = 128;  // ray indices a wave claims from the queue at once
// One-ray-per-lane kernel: stack entries per lane in LDS.  With 16 (16 KB
// per 256-thread block) LDS allows 8 waves per SIMD and deeper entries, up to
// kStackDepth, go to a per-lane global spill area; with kStackDepth nothing
// spills but LDS allows 5 waves per SIMD.  The launcher takes the full LDS
// stack when its whole grid is resident at once (a small batch gains nothing
// from more waves and would pay for the spills), the short one otherwise.
constexpr int kSimpleLds = 16;
// Speculative any-hit leaves the node loop for the leaf phase once this many
// lanes or fewer still look for a leaf (and some lane holds one).
constexpr uint32_t kLeafPhaseAt = 8;

struct HitOutFull {  // CRT_Hit (CrossRT.h:20-27)
  float t;
  uint32_t prim, inst, geom;
  float coords[4];
};

__global__ void k_reset_counter(uint32_t* c) { *c = 0u; }

// Reference-order reference kernel (one ray per lane, no refill).  Kept for
// the counting variant used to report the reference's traversal counters.
template <bool kForceAny, bool kStats, int kLds, class Tri = MtTest>
__global__ __launch_bounds__(256) void k_trace_simple(const NodePair* __restrict__ pairs,
                                                      const TriPre* __restrict__ tris,
                                                      const Ray* __restrict__ rays, uint64_t n, void* out,
                                                      int format, uint2 ids, uint32_t* __restrict__ stats,
                                                      uint32_t* __restrict__ spill) {
  __shared__ uint32_t stack_mem[kLds * 256];
  const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const ShortStack<kLds, 256> stk{stack_mem + threadIdx.x, spill + i, gridDim.x * 256u};
  const Ray r = rays[i];
  const float3 o = make_float3(r.pos_near.x, r.pos_near.y, r.pos_near.z);
  const float3 d = make_float3(r.dir_far.x, r.dir_far.y, r.dir_far.z);
  const float tnear = r.pos_near.w;
  const bool any = kForceAny || (r.dir_far.w <= 0.0f);
  const float tfar = kForceAny ? fabsf(r.dir_far.w) : (any ? -r.dir_far.w : r.dir_far.w);
  const float3 inv = safe_inverse(d);
  const Tri tt(o, d);
  HitRec h;
  h.t = tfar;
  h.prim = kInvalid;
  h.u = 0.0f;
  h.v = 0.0f;
  uint32_t npairs = 0, ntests = 0;
  int top = 0;
  uint32_t cur = 0;
  while (true) {
    bool pop;
    if ((cur & kLeafBit) == 0) {
      const float4* pp = reinterpret_cast<const float4*>(pairs + cur);
      const float4 lmin = pp[0], lmax = pp[1], rmin = pp[2], rmax = pp[3];
      if (kStats) npairs++;
      const float2 t0 = tt.box(o, inv, lmin.x, lmin.y, lmin.z, lmax.x, lmax.y, lmax.z);
      const float2 t1 = tt.box(o, inv, rmin.x, rmin.y, rmin.z, rmax.x, rmax.y, rmax.z);
      const bool h0 = (t0.x <= t0.y) && (t0.y >= tnear) && (t0.x <= h.t);
      const bool h1 = (t1.x <= t1.y) && (t1.y >= tnear) && (t1.x <= h.t);
      const uint32_t c0 = __float_as_uint(lmin.w), c1 = __float_as_uint(rmin.w);
      pop = !h0 && !h1;
      cur = h0 ? c0 : c1;
      if (h0 && h1) {
        const bool left_first = t0.x <= t1.x;
        cur = left_first ? c0 : c1;
        stk.push(top, left_first ? c1 : c0);
      }
    } else {
      pop = true;
      if (cur != kInvalid) {
        const uint32_t s = leaf_start(cur), c = leaf_count(cur);
        if (kStats) ntests += c;
        for (uint32_t k = s; k < s + c; k++) tri_test(tt, o, d, tnear, tris[k], h);
        if (any && h.prim != kInvalid) break;
      }
    }
    if (pop) {
      if (top == 0) break;
      cur = stk.pop(top);
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
    o4[0] = make_float4(h.t, __uint_as_float(h.prim), __uint_as_float(hit ? ids.x : kInvalid),
                        __uint_as_float(hit ? ids.y : kInvalid));
    o4[1] = make_float4(h.u, h.v, 0.0f, 0.0f);
  }
}

// Persistent wave64 traversal with ballot-based refill.
//   kForceAny: every ray is an any-hit query (hrt_trace_any)
//   kSpec:     speculative while-while traversal (any-hit only)
template <bool kForceAny, bool kSpec, class Tri = MtTest>
__global__ __launch_bounds__(kTraceBlock) void k_trace_persistent(const NodePair* __restrict__ pairs,
                                                                  const TriPre* __restrict__ tris,
                                                                  const Ray* __restrict__ rays, uint64_t n,
                                                                  void* out, int format, uint2 ids,
                                                                  uint32_t* __restrict__ queue,
                                                                  uint32_t* __restrict__ spill) {
  __shared__ uint32_t stack_mem[kShortStack * kTraceBlock];
  const uint32_t lanes = gridDim.x * kTraceBlock;
  const ShortStack<kShortStack, kTraceBlock> stk{stack_mem + threadIdx.x,
                                                 spill + blockIdx.x * kTraceBlock + threadIdx.x, lanes};

  // per-lane ray state
  bool active = false;
  uint64_t rid = 0;
  float3 o = make_float3(0.f, 0.f, 0.f), d = o, inv = o;
  float tnear = 0.f;
  bool any = kForceAny;
  Tri tt;
  HitRec h;
  h.t = 0.f;
  h.prim = kInvalid;
  h.u = h.v = 0.f;
  int top = 0;          // entries on the stack
  uint32_t cur = kInvalid;   // node pair to test next, a leaf, or kInvalid
  uint32_t leaf = 0;    // parked leaf (speculative any-hit), 0 = none
  RayPool pool;  // ray indices claimed by this wave

  while (true) {
    // ---- refill idle lanes from the global queue -------------------------
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = __popcll(idle);
    if (!pool.drained() && (nidle >= kRefillAt || nidle == kWave)) {
      const uint64_t idx = pool.take<kClaim>(nidle, lanes_below(idle), queue, n);
      if (!active) {
        if (idx < n) {
          rid = idx;
          const Ray r = rays[idx];
          o = make_float3(r.pos_near.x, r.pos_near.y, r.pos_near.z);
          d = make_float3(r.dir_far.x, r.dir_far.y, r.dir_far.z);
          tnear = r.pos_near.w;
          any = kForceAny || (r.dir_far.w <= 0.0f);
          h.t = kForceAny ? fabsf(r.dir_far.w) : (any ? -r.dir_far.w : r.dir_far.w);
          h.prim = kInvalid;
          h.u = h.v = 0.f;
          inv = safe_inverse(d);
          tt = Tri(o, d);
          top = 0;
          cur = 0;
          leaf = 0;
          active = true;
        }
      }
    }
    if (__ballot(active) == 0) break;

    // ---- traverse until enough lanes are idle -----------------------------
    while (true) {
      bool done = false;
      if (kSpec) {
        // inner phase: lanes without a parked leaf search for one; lanes
        // holding one keep descending while their next node is internal
        while (true) {
          // park a leaf reached in `cur` (at most one per lane)
          if (active && leaf == 0 && (cur & kLeafBit) != 0 && cur != kInvalid) {
            leaf = cur;
            cur = (top > 0) ? stk.pop(top) : kInvalid;
          }
          const bool searching = active && leaf == 0 && (cur & kLeafBit) == 0;
          const uint32_t nsearch = __popcll(__ballot(searching));
          if (nsearch == 0 || (nsearch <= kLeafPhaseAt && __any(leaf != 0))) break;
          if (active && (cur & kLeafBit) == 0) {
            const float4* pp = reinterpret_cast<const float4*>(pairs + cur);
            const float4 lmin = pp[0], lmax = pp[1], rmin = pp[2], rmax = pp[3];
            const float2 t0 = tt.box(o, inv, lmin.x, lmin.y, lmin.z, lmax.x, lmax.y, lmax.z);
            const float2 t1 = tt.box(o, inv, rmin.x, rmin.y, rmin.z, rmax.x, rmax.y, rmax.z);
            const bool h0 = (t0.x <= t0.y) && (t0.y >= tnear) && (t0.x <= h.t);
            const bool h1 = (t1.x <= t1.y) && (t1.y >= tnear) && (t1.x <= h.t);
            const uint32_t c0 = __float_as_uint(lmin.w), c1 = __float_as_uint(rmin.w);
            if (h0 && h1) {
              const bool left_first = t0.x <= t1.x;
              cur = left_first ? c0 : c1;
              stk.push(top, left_first ? c1 : c0);
            } else if (h0 || h1) {
              cur = h0 ? c0 : c1;
            } else {
              cur = (top > 0) ? stk.pop(top) : kInvalid;
            }
          }
        }
        // leaf phase
        if (active && leaf != 0) {
          const uint32_t s = leaf_start(leaf), c = leaf_count(leaf);
          for (uint32_t k = s; k < s + c; k++) {
            if (tri_hit_before(tt, o, d, tnear, tris[k], h.t)) {  // any-hit: the first hit decides
              h.prim = 0u;
              break;
            }
          }
          leaf = 0;
        }
        done = active && (h.prim != kInvalid || (leaf == 0 && cur == kInvalid));
      } else {
        // one step in the reference's order: a pair or a leaf
        if (active) {
          if ((cur & kLeafBit) == 0) {
            const float4* pp = reinterpret_cast<const float4*>(pairs + cur);
            const float4 lmin = pp[0], lmax = pp[1], rmin = pp[2], rmax = pp[3];
            const float2 t0 = tt.box(o, inv, lmin.x, lmin.y, lmin.z, lmax.x, lmax.y, lmax.z);
            const float2 t1 = tt.box(o, inv, rmin.x, rmin.y, rmin.z, rmax.x, rmax.y, rmax.z);
            const bool h0 = (t0.x <= t0.y) && (t0.y >= tnear) && (t0.x <= h.t);
            const bool h1 = (t1.x <= t1.y) && (t1.y >= tnear) && (t1.x <= h.t);
            const uint32_t c0 = __float_as_uint(lmin.w), c1 = __float_as_uint(rmin.w);
            if (h0 && h1) {
              const bool left_first = t0.x <= t1.x;
              cur = left_first ? c0 : c1;
              stk.push(top, left_first ? c1 : c0);
            } else if (h0 || h1) {
              cur = h0 ? c0 : c1;
            } else {
              cur = (top > 0) ? stk.pop(top) : kInvalid;
            }
          } else if (cur != kInvalid) {
            const uint32_t s = leaf_start(cur), c = leaf_count(cur);
            for (uint32_t k = s; k < s + c; k++) tri_test(tt, o, d, tnear, tris[k], h);
            cur = (top > 0) ? stk.pop(top) : kInvalid;
          }
          done = (cur == kInvalid) || (any && h.prim != kInvalid);
        }
      }
      if (done) {
        if (kForceAny) {
          static_cast<uint8_t*>(out)[rid] = (h.prim != kInvalid) ? 1 : 0;
        } else if (format == kHitCompact) {
          static_cast<uint2*>(out)[rid] = make_uint2(__float_as_uint(h.t), h.prim);
        } else {
          const bool hit = h.prim != kInvalid;
          float4* o4 = reinterpret_cast<float4*>(static_cast<HitOutFull*>(out) + rid);
          o4[0] = make_float4(h.t, __uint_as_float(h.prim), __uint_as_float(hit ? ids.x : kInvalid),
                              __uint_as_float(hit ? ids.y : kInvalid));
          o4[1] = make_float4(h.u, h.v, 0.0f, 0.0f);
        }
        active = false;
      }
      const uint32_t nact = __popcll(__ballot(active));
      if (nact == 0 || (!pool.drained() && kWave - nact >= kRefillAt)) break;
    }
  }
}

template <bool kAny, bool kSpec, class Tri>
hipError_t launch_persistent(const Blas& b, const TraceArgs& a, hipStream_t stream) {
  const void* kfn = reinterpret_cast<const void*>(&k_trace_persistent<kAny, kSpec, Tri>);
  uint64_t want = (a.n + kTraceBlock - 1) / kTraceBlock;
  constexpr size_t kLaneWords = kStackDepth - kShortStack;
  uint64_t cap = resident_grid(kfn, kTraceBlock);
  if (a.workspace) cap = std::min<uint64_t>(cap, (a.workspace_words - 64) / (kTraceBlock * kLaneWords));
  const int grid = int(std::max<uint64_t>(1, std::min<uint64_t>(want, cap)));
  // queue counter + spill area for stack entries beyond the LDS part
  const size_t spill_words = size_t(grid) * kTraceBlock * kLaneWords;
  uint32_t* buf = a.workspace;
  hipError_t e = hipSuccess;
  if (!buf && (e = hipMallocAsync(reinterpret_cast<void**>(&buf), sizeof(uint32_t) * (64 + spill_words), stream)) !=
                  hipSuccess)
    return e;
  k_reset_counter<<<1, 1, 0, stream>>>(buf);
  k_trace_persistent<kAny, kSpec, Tri><<<grid, kTraceBlock, 0, stream>>>(
      b.pairs, tri_records<Tri>(b), a.rays, a.n, a.out, a.format, make_uint2(a.inst_id, a.geom_id), buf, buf + 64);
  e = hipGetLastError();
  hipError_t f = a.workspace ? hipSuccess : hipFreeAsync(buf, stream);
  return e != hipSuccess ? e : f;
}

template <int kLds, class Tri>
hipError_t launch_simple(const Blas& b, const TraceArgs& a, uint32_t blocks, uint32_t* spill, hipStream_t stream) {
  const uint2 ids = make_uint2(a.inst_id, a.geom_id);
  const TriPre* tris = tri_records<Tri>(b);
  if (a.stats && a.anyhit)
    k_trace_simple<true, true, kLds, Tri><<<blocks, 256, 0, stream>>>(b.pairs, tris, a.rays, a.n, a.out, a.format,
                                                                       ids, a.stats, spill);
  else if (a.stats)
    k_trace_simple<false, true, kLds, Tri><<<blocks, 256, 0, stream>>>(b.pairs, tris, a.rays, a.n, a.out, a.format,
                                                                        ids, a.stats, spill);
  else
    k_trace_simple<false, false, kLds, Tri><<<blocks, 256, 0, stream>>>(b.pairs, tris, a.rays, a.n, a.out,
                                                                         a.format, ids, nullptr, spill);
  return hipGetLastError();
}

template <class Tri>
hipError_t trace_blas_with(const Blas& b, const TraceArgs& a, hipStream_t stream) {
  if (a.anyhit && !a.stats) return launch_persistent<true, true, Tri>(b, a, stream);
  // One ray per lane: the counting variant (the reference's order) and
  // nearest hits (rays that arrive in screen order stay coherent within a
  // wave, which the refilling kernel gives up).
  const uint32_t blocks = uint32_t((a.n + 255) / 256);
  if (blocks <= resident_grid(reinterpret_cast<const void*>(&k_trace_simple<false, false, kStackDepth, Tri>), 256))
    return launch_simple<kStackDepth, Tri>(b, a, blocks, nullptr, stream);
  uint32_t* spill = nullptr;
  hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&spill),
                                sizeof(uint32_t) * size_t(blocks) * 256 * (kStackDepth - kSimpleLds), stream);
  if (e != hipSuccess) return e;
  e = launch_simple<kSimpleLds, Tri>(b, a, blocks, spill, stream);
  const hipError_t f = hipFreeAsync(spill, stream);
  return e != hipSuccess ? e : f;
}

}  // namespace

void warm_bvh2_any_hit_launchers() {
  (void)resident_grid(reinterpret_cast<const void*>(&k_trace_persistent<true, true, MtTest>), kTraceBlock);
  (void)resident_grid(reinterpret_cast<const void*>(&k_trace_persistent<true, true, WtTest>), kTraceBlock);
}

uint32_t resident_grid(const void* kernel, int block) {
  struct Entry {
    int dev;
    const void* kernel;
    int block;
    uint32_t grid;
  };
  static std::mutex mu;
  static std::vector<Entry> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  {
    std::lock_guard<std::mutex> lock(mu);
    for (const Entry& e : cache)
      if (e.dev == dev && e.kernel == kernel && e.block == block) return e.grid;
  }
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu <= 0)
    per_cu = 1;
  const uint32_t grid = uint32_t(cus) * uint32_t(per_cu);
  std::lock_guard<std::mutex> lock(mu);
  cache.push_back(Entry{dev, kernel, block, grid});
  return grid;
}

hipError_t trace_blas(const Blas& b, const TraceArgs& a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  if (a.watertight) {
    if (!b.tris_wt) return hipErrorInvalidValue;
    return trace_blas_with<WtTest>(b, a, stream);
  }
  return trace_blas_with<MtTest>(b, a, stream);
}

}  // namespace hrt
