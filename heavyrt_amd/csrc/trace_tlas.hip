// trace_tlas.hip -- two-level (instanced) nearest-hit / any-hit traversal.
//
// Semantics follow BVH2CommonRT::RayQuery_NearestHit (core/BVH2CommonRT.cpp:
// 228-279): the ray is tested against the top-level tree over instance boxes;
// for every instance it reaches, the ray is moved into object space with the
// instance's inverse matrix (matmul4x3 / matmul3x3, raytrace_common.h:82-90;
// the direction is not renormalised, so t stays a world-space distance) and
// the object's BVH2 is walked exactly as the single-instance kernel walks it,
// carrying the running hit between instances.  Two deliberate differences
// from the reference:
//   * the reference collects instance leaves first (at most LBVH_MAXHITS =
//     32 of them, BVH2CommonRT.cpp:213-221) and uses a placeholder entry
//     distance; here the walk is interleaved, culls with the running hit
//     distance and has no limit on the number of instances a ray meets;
//   * on equal t across instances the smaller instance id wins, so the
//     answer does not depend on the shape of the top-level tree.
//
// Layout: the top-level tree is NodePair pairs over instance boxes (the same
// builder as the BLAS, fed with boxes), leaves index an InstDev array in leaf
// order (80 B each: inverse rows, the object's pairs and triangles, ids).
//
// Execution: a persistent wave64 kernel with ballot-based refill, one node
// pair or one leaf per lane per step, and a per-lane stack of 64 entries: 16
// in LDS, the rest in a per-lane global spill area.  64 covers the deepest
// possible walk: at most 30 pending entries in the top-level tree plus 30 in
// one object's tree (Karras trees over unique 30-bit codes have at most 30
// internal levels).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "hrt_internal.h"

namespace hrt {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = kWave;
constexpr int kRefillAt = 16;
constexpr uint32_t kClaim = 128;  // ray indices a wave claims from the queue at once
constexpr int kTwoLevelDepth = 64;

__global__ void k_reset_counter(uint32_t* c) { *c = 0u; }

// One step of the near-first pair test shared by both levels: returns the next
// node to visit, pushes the deferred child, or reports that both missed.
__device__ inline bool pair_step(const NodePair* __restrict__ P, uint32_t& cur, float3 o, float3 inv, float tnear,
                                 float tmax, const ShortStack<kShortStack, kBlock>& stk, int& top) {
  const float4* pp = reinterpret_cast<const float4*>(P + cur);
  const float4 lmin = pp[0], lmax = pp[1], rmin = pp[2], rmax = pp[3];
  const float2 t0 = slab(o, inv, lmin.x, lmin.y, lmin.z, lmax.x, lmax.y, lmax.z);
  const float2 t1 = slab(o, inv, rmin.x, rmin.y, rmin.z, rmax.x, rmax.y, rmax.z);
  const bool h0 = (t0.x <= t0.y) && (t0.y >= tnear) && (t0.x <= tmax);
  const bool h1 = (t1.x <= t1.y) && (t1.y >= tnear) && (t1.x <= tmax);
  const uint32_t c0 = __float_as_uint(lmin.w), c1 = __float_as_uint(rmin.w);
  if (h0 && h1) {
    const bool left_first = t0.x <= t1.x;
    cur = left_first ? c0 : c1;
    stk.push(top, left_first ? c1 : c0);
    return true;
  }
  if (h0 || h1) {
    cur = h0 ? c0 : c1;
    return true;
  }
  return false;
}

template <bool kForceAny, bool kStats>
__global__ __launch_bounds__(kBlock) void k_trace_two_level(const NodePair* __restrict__ tlas,
                                                            const InstDev* __restrict__ inst,
                                                            const Ray* __restrict__ rays, uint64_t n, void* out,
                                                            int format, uint32_t* __restrict__ stats,
                                                            uint32_t* __restrict__ queue,
                                                            uint32_t* __restrict__ spill) {
  __shared__ uint32_t stack_mem[kShortStack * kBlock];
  const uint32_t lanes = gridDim.x * kBlock;
  const ShortStack<kShortStack, kBlock> stk{stack_mem + threadIdx.x, spill + blockIdx.x * kBlock + threadIdx.x,
                                             lanes};

  bool active = false;
  uint64_t rid = 0;
  float3 wo = make_float3(0.f, 0.f, 0.f), wd = wo, winv = wo;  // world-space ray
  float3 o = wo, d = wo, inv = wo;                               // ray in the current tree's space
  float tnear = 0.f;
  bool any = kForceAny;
  HitRec h;
  h.t = 0.f;
  h.prim = kInvalid;
  h.u = h.v = 0.f;
  uint32_t h_inst = kInvalid, h_geom = kInvalid;
  const NodePair* P = tlas;  // tree being walked
  const TriPre* T = nullptr;
  bool in_blas = false;
  int top = 0, inst_top = 0;
  uint32_t k = 0, kend = 0;  // instances of the current top-level leaf
  uint32_t cur_inst = 0, cur_geom = 0;
  uint32_t cur = kInvalid;
  uint32_t npairs = 0, ntests = 0;
  RayPool pool;  // ray indices claimed by this wave

  auto enter = [&](uint32_t slot) {
    const InstDev* I = inst + slot;
    const float4 r0 = I->inv_row[0], r1 = I->inv_row[1], r2 = I->inv_row[2];
    o = xform_point(r0, r1, r2, wo);
    d = xform_dir(r0, r1, r2, wd);
    inv = safe_inverse(d);
    P = I->pairs;
    T = I->tris;
    cur_inst = I->inst;
    cur_geom = I->geom;
    in_blas = true;
    inst_top = top;
    cur = 0;
  };
  // the current subtree is done: next instance of the leaf, or pop
  auto next = [&]() {
    if (in_blas && top == inst_top) {
      if (++k < kend) {
        enter(k);
        return;
      }
      in_blas = false;
      P = tlas;
      o = wo;
      d = wd;
      inv = winv;
    }
    cur = (top > 0) ? stk.pop(top) : kInvalid;
  };

  while (true) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = __popcll(idle);
    if (!pool.drained() && (nidle >= kRefillAt || nidle == kWave)) {
      const uint64_t idx = pool.take<kClaim>(nidle, lanes_below(idle), queue, n);
      if (!active) {
        if (idx < n) {
          rid = idx;
          const Ray r = rays[idx];
          wo = make_float3(r.pos_near.x, r.pos_near.y, r.pos_near.z);
          wd = make_float3(r.dir_far.x, r.dir_far.y, r.dir_far.z);
          winv = safe_inverse(wd);
          o = wo;
          d = wd;
          inv = winv;
          tnear = r.pos_near.w;
          any = kForceAny || (r.dir_far.w <= 0.0f);
          h.t = kForceAny ? fabsf(r.dir_far.w) : (any ? -r.dir_far.w : r.dir_far.w);
          h.prim = kInvalid;
          h.u = h.v = 0.f;
          h_inst = h_geom = kInvalid;
          P = tlas;
          in_blas = false;
          top = 0;
          inst_top = 0;
          k = kend = 0;
          cur = 0;
          npairs = ntests = 0;
          active = true;
        }
      }
    }
    if (__ballot(active) == 0) break;

    while (true) {
      bool done = false;
      if (active) {
        if ((cur & kLeafBit) == 0) {
          if (kStats) npairs++;
          if (!pair_step(P, cur, o, inv, tnear, h.t, stk, top)) next();
        } else if (in_blas) {
          const uint32_t s = leaf_start(cur), c = leaf_count(cur);
          if (kStats) ntests += c;
          for (uint32_t j = s; j < s + c; j++)
            if (mt_test_inst(o, d, tnear, T[j], cur_inst, h, h_inst)) h_geom = cur_geom;
          next();
        } else {
          const uint32_t s = leaf_start(cur), c = leaf_count(cur);
          if (c > 0) {
            k = s;
            kend = s + c;
            enter(k);
          } else {
            next();
          }
        }
        done = (cur == kInvalid) || (any && h.prim != kInvalid);
      }
      if (done) {
        const bool hit = h.prim != kInvalid;
        if (kStats) {
          stats[2 * rid + 0] = npairs;
          stats[2 * rid + 1] = ntests;
        }
        if (kForceAny) {
          static_cast<uint8_t*>(out)[rid] = hit ? 1 : 0;
        } else if (format == kHitCompact) {
          static_cast<uint2*>(out)[rid] = make_uint2(__float_as_uint(h.t), h.prim);
        } else {
          float4* o4 = reinterpret_cast<float4*>(static_cast<char*>(out) + 32 * rid);
          o4[0] = make_float4(h.t, __uint_as_float(h.prim), __uint_as_float(hit ? h_inst : kInvalid),
                              __uint_as_float(hit ? h_geom : kInvalid));
          o4[1] = make_float4(h.u, h.v, 0.0f, 0.0f);
        }
        active = false;
      }
      const uint32_t nact = __popcll(__ballot(active));
      if (nact == 0 || (!pool.drained() && kWave - nact >= kRefillAt)) break;
    }
  }
}

// Misses for every ray of a scene without instances.
__global__ __launch_bounds__(256) void k_misses(const Ray* __restrict__ rays, uint64_t n, void* out, int format,
                                                int force_any, uint32_t* __restrict__ stats) {
  const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  if (stats) {
    stats[2 * i + 0] = 0u;
    stats[2 * i + 1] = 0u;
  }
  if (force_any) {
    static_cast<uint8_t*>(out)[i] = 0;
    return;
  }
  const float w = rays[i].dir_far.w;
  const float t = (w <= 0.0f) ? -w : w;
  if (format == kHitCompact) {
    static_cast<uint2*>(out)[i] = make_uint2(__float_as_uint(t), kInvalid);
  } else {
    float4* o4 = reinterpret_cast<float4*>(static_cast<char*>(out) + 32 * i);
    o4[0] = make_float4(t, __uint_as_float(kInvalid), __uint_as_float(kInvalid), __uint_as_float(kInvalid));
    o4[1] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

int resident_grid(const void* kernel) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || per_cu <= 0)
    per_cu = 8;
  return cus * per_cu;
}

template <bool kAny, bool kStats>
hipError_t launch(const Blas& tlas, const InstDev* inst, const TraceArgs& a, hipStream_t stream) {
  const void* kfn = reinterpret_cast<const void*>(&k_trace_two_level<kAny, kStats>);
  const uint64_t want = (a.n + kBlock - 1) / kBlock;
  const int grid = int(std::min<uint64_t>(want, uint64_t(resident_grid(kfn))));
  const size_t spill_words = size_t(grid) * kBlock * (kTwoLevelDepth - kShortStack);
  uint32_t* buf = nullptr;
  hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&buf), sizeof(uint32_t) * (64 + spill_words), stream);
  if (e != hipSuccess) return e;
  k_reset_counter<<<1, 1, 0, stream>>>(buf);
  k_trace_two_level<kAny, kStats><<<grid, kBlock, 0, stream>>>(tlas.pairs, inst, a.rays, a.n, a.out, a.format,
                                                               a.stats, buf, buf + 64);
  e = hipGetLastError();
  hipError_t f = hipFreeAsync(buf, stream);
  return e != hipSuccess ? e : f;
}

}  // namespace

hipError_t trace_two_level(const Blas& tlas, const InstDev* inst, const TraceArgs& a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  if (tlas.ntris == 0 || !inst) {
    const uint32_t blocks = uint32_t((a.n + 255) / 256);
    k_misses<<<blocks, 256, 0, stream>>>(a.rays, a.n, a.out, a.format, a.anyhit, a.stats);
    return hipGetLastError();
  }
  if (a.anyhit) return a.stats ? launch<true, true>(tlas, inst, a, stream) : launch<true, false>(tlas, inst, a, stream);
  return a.stats ? launch<false, true>(tlas, inst, a, stream) : launch<false, false>(tlas, inst, a, stream);
}

}  // namespace hrt
