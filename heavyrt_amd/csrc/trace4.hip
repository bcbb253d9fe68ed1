// trace4.hip -- BVH4 nearest-hit / any-hit traversal (FP32 or FP16 boxes).
//
// Visiting order follows BVH4HalfRT::BVH4Traverse (core/BVH4HalfRT.cpp:73-256):
// the four child boxes are slab-tested in FP32 (FP16 boxes are widened to FP32
// first, UnpackBVHNode :56-71), a missed child gets MAXFLOAT, a five-comparator
// network sorts the (distance, child) pairs (:119-180), the hit children w, z,
// y are pushed and the walk continues with x (:184-222).  Leaves are tested
// with the reference's Moller-Trumbore; the nearest hit is chosen by
// (t, primId), so FP16 and FP32 nodes -- which visit in different orders --
// return identical hits.  tFar <= 0 or hrt_trace_any: stop at the first hit.
//
// The stack holds at most 45 entries on a tree over unique 30-bit codes (BVH4
// nodes are the even-depth nodes of a BVH2 with at most 30 internal levels:
// 15 node levels, 3 deferred children each) and at most kStackDepth4 = 63 on
// the clustered second tree, whose build keeps it within 42 levels.  The
// persistent kernels keep 16 entries per lane in LDS and the one-ray-per-lane
// kernel 48, interleaved by lane; deeper entries go to a per-lane global
// spill area, so LDS does not limit occupancy.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ao.h"
#include "hrt_internal.h"

namespace hrt {
namespace {

constexpr int kWave = 64;
constexpr int kBlock4 = 64;
constexpr int kRefillAt4 = 40;
// LDS stack entries per lane of the one-ray-per-lane kernels (the rest spill
// to global memory).  Deep, incoherent nearest-hit walks push often and
// measured slower with spills, so BVH4 keeps its whole stack in LDS.
constexpr int kSimpleLds4 = 48;
// Nearest-hit batches of at least this many rays are sorted into coherent
// waves before they are traced.  On 4,194,304 incoherent rays in a
// 4,194,304-triangle soup the FP16 walk fetches 6 % fewer bytes from beyond
// the L2 and takes 21.8 instead of 24.0 ms; a refilling persistent kernel
// (wave-local ray pool, speculative while-while) was no faster there (24.8
// ms unsorted, 23.3 sorted): the walk waits on memory, not on idle lanes.
constexpr uint64_t kSortRaysMin4 = 1u << 20;
// The speculative any-hit kernels leave the node loop for the leaf phase once
// this many lanes or fewer still look for a leaf (0 would wait for the last).
constexpr uint32_t kLeafPhaseAt = 4;
constexpr uint32_t kClaim = 128;  // ray indices a wave claims from the queue at once
constexpr float kMaxFloat = 3.402823466e+38f;

struct HitOutFull {
  float t;
  uint32_t prim, inst, geom;
  float coords[4];
};

__global__ void k_reset_counter4(uint32_t* c) { *c = 0u; }

// Loads one BVH4 node and tests its four children with the box test of the
// triangle-test policy tt: bt[k] = the child's (entry, exit) distances,
// ch[k] its reference.  kFused (FP16 nodes of any-hit traversals under the
// reference's test only): slab_fused with oi = -(o * inv) for a lane whose
// fz = fused_ok(oi), slab otherwise.
template <bool kFp16, bool kFused, class Tri>
__device__ inline void quad_boxes(const void* __restrict__ quads, uint32_t q, float3 o, float3 inv, float3 oi,
                                  bool fz, const Tri& tt, float2 bt[4], uint32_t ch[4]) {
  static_assert(kFp16 || !kFused, "the fused box test is for FP16 nodes");
  if (kFp16) {
    // per child {minx | miny << 16, minz | maxx << 16, maxy | maxz << 16, ref};
    // the FP16 values are widened inside each branch, so the fused one feeds
    // them straight to v_fma_mix_f32
    const uint4* p = reinterpret_cast<const uint4*>(static_cast<const Quad16*>(quads) + q);
    uint4 w[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      w[k] = p[k];
      ch[k] = w[k].w;
    }
    if (kFused && !Tri::kRobustBoxes && fz) {
      const uint3 sel = near_far_selectors(inv);
#pragma unroll
      for (int k = 0; k < 4; k++) bt[k] = slab_fused(oi, inv, w[k], sel);
    } else {
#pragma unroll
      for (int k = 0; k < 4; k++)
        bt[k] = tt.box(o, inv, h2f_lo(w[k].x), h2f_hi(w[k].x), h2f_lo(w[k].y), h2f_hi(w[k].y), h2f_lo(w[k].z),
                       h2f_hi(w[k].z));
    }
  } else {
    const float4* p = reinterpret_cast<const float4*>(static_cast<const Quad32*>(quads) + q);
    const float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4], f = p[5];
    const uint4 r = reinterpret_cast<const uint4*>(p)[6];
    const float lox[4] = {a.x, a.y, a.z, a.w}, hix[4] = {b.x, b.y, b.z, b.w};
    const float loy[4] = {c.x, c.y, c.z, c.w}, hiy[4] = {d.x, d.y, d.z, d.w};
    const float loz[4] = {e.x, e.y, e.z, e.w}, hiz[4] = {f.x, f.y, f.z, f.w};
    ch[0] = r.x; ch[1] = r.y; ch[2] = r.z; ch[3] = r.w;
#pragma unroll
    for (int k = 0; k < 4; k++) bt[k] = tt.box(o, inv, lox[k], loy[k], loz[k], hix[k], hiy[k], hiz[k]);
  }
}

// Whether a child with box distances t and reference c is entered.  An unused
// slot (box +1e9..-1e9, empty leaf 0x80000000) would pass the min/max slab
// test as the box [-1e9, 1e9]; kSkipEmpty keeps it out.  slab_fused picks
// each axis's near and far plane by the sign of the ray's direction, so it
// finds the inverted box empty by itself; a caller that mostly uses it may
// leave kSkipEmpty off, and a lane on the min/max form then enters the empty
// leaf, which tests no triangle.
template <bool kSkipEmpty = true>
__device__ inline bool child_hit(float2 t, uint32_t c, float tnear, float tcull) {
  return (t.x <= t.y) && (t.y >= tnear) && (t.x <= tcull) && (!kSkipEmpty || c != kEmptyLeaf);
}

// quad_boxes, then the entry distances (kMaxFloat for a miss) sorted
// ascending with their references, as the reference orders a node's children.
template <bool kFp16, bool kFused, class Tri>
__device__ inline void visit_quad(const void* __restrict__ quads, uint32_t q, float3 o, float3 inv, float3 oi,
                                  bool fz, const Tri& tt, float tnear, float tcull, float tm[4], uint32_t ch[4]) {
  float2 bt[4];
  quad_boxes<kFp16, kFused>(quads, q, o, inv, oi, fz, tt, bt, ch);
#pragma unroll
  for (int k = 0; k < 4; k++) tm[k] = child_hit(bt[k], ch[k], tnear, tcull) ? bt[k].x : kMaxFloat;
  // five-comparator network, BVH4HalfRT.cpp:121-180 (swap when strictly less)
  auto cswap = [&](int i, int j) {
    const bool s = tm[j] < tm[i];
    const float ti = tm[i], tj = tm[j];
    const uint32_t ci = ch[i], cj = ch[j];
    tm[i] = s ? tj : ti;
    tm[j] = s ? ti : tj;
    ch[i] = s ? cj : ci;
    ch[j] = s ? ci : cj;
  };
  cswap(0, 1);
  cswap(2, 3);
  cswap(0, 2);
  cswap(1, 3);
  cswap(1, 2);
}

template <bool kForceAny>
__device__ inline void write_hit(void* out, uint64_t i, int format, uint2 ids, const HitRec& h) {
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

// One ray per lane, the reference's order step by step (also the counting
// variant: quads visited, triangles in visited leaves).
template <bool kForceAny, bool kFp16, bool kStats, class Tri = MtTest>
__global__ __launch_bounds__(kBlock4) void k_trace4_simple(const void* __restrict__ quads,
                                                           const TriPre* __restrict__ tris,
                                                           const Ray* __restrict__ rays, uint64_t n, void* out,
                                                           int format, uint2 ids, uint32_t* __restrict__ stats,
                                                           uint32_t* __restrict__ spill,
                                                           const uint32_t* const* __restrict__ order_slot) {
  __shared__ uint32_t stack_mem[kSimpleLds4 * kBlock4];
  const uint32_t* __restrict__ order = order_slot ? *order_slot : nullptr;
  // a sorted batch: each XCD takes one contiguous eighth of the sorted order,
  // so its L2 serves rays of similar direction and origin (5 % faster on 4 M
  // incoherent rays); rays in their own order keep the dispatch order, whose
  // interleaving balances the XCDs over uneven screen regions
  const uint64_t lane_id =
      uint64_t(order ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x) * kBlock4 + threadIdx.x;
  if (lane_id >= n) return;
  // thread k traces ray order[k] when the batch was sorted into coherent waves
  const uint64_t i = order ? uint64_t(order[lane_id]) : lane_id;
  if (i >= n) return;
  const ShortStack<kSimpleLds4, kBlock4> stk{stack_mem + threadIdx.x, spill + lane_id, gridDim.x * kBlock4};
  const Ray r = rays[i];
  const float3 o = make_float3(r.pos_near.x, r.pos_near.y, r.pos_near.z);
  const float3 d = make_float3(r.dir_far.x, r.dir_far.y, r.dir_far.z);
  const float tnear = r.pos_near.w;
  const bool any = kForceAny || (r.dir_far.w <= 0.0f);
  const float3 inv = safe_inverse(d);
  const float3 oi = neg_o_inv(o, inv);
  const bool fz = fused_ok(oi);
  const Tri tt(o, d);
  HitRec h;
  h.t = kForceAny ? fabsf(r.dir_far.w) : (any ? -r.dir_far.w : r.dir_far.w);
  h.prim = kInvalid;
  h.u = h.v = 0.0f;
  uint32_t nvisits = 0, ntests = 0;
  int top = 0;
  uint32_t cur = 0;
  while (true) {
    bool pop = false;
    if ((cur & kLeafBit) == 0) {
      float tm[4];
      uint32_t ch[4];
      visit_quad<kFp16, kForceAny && kFp16>(quads, cur, o, inv, oi, fz, tt, tnear, h.t, tm, ch);
      if (kStats) nvisits++;
      if (tm[3] < kMaxFloat) stk.push(top, ch[3]);
      if (tm[2] < kMaxFloat) stk.push(top, ch[2]);
      if (tm[1] < kMaxFloat) stk.push(top, ch[1]);
      if (tm[0] < kMaxFloat) cur = ch[0];
      else pop = true;
    }
    if (!pop && (cur & kLeafBit) != 0 && cur != kInvalid) {
      const uint32_t s = leaf_start(cur), c = leaf_count(cur);
      if (kStats) ntests += c;
      for (uint32_t k = s; k < s + c; k++)
        if (tri_test_lex(tt, o, d, tnear, tris[k], h) && any) break;
      pop = true;
      if (any && h.prim != kInvalid) break;
    }
    if (pop) {
      if (top == 0) break;
      cur = stk.pop(top);
    }
  }
  if (kStats) {
    stats[2 * i + 0] = nvisits;
    stats[2 * i + 1] = ntests;
  }
  write_hit<kForceAny>(out, i, format, ids, h);
}

// Persistent speculative any-hit over BVH4 nodes with ballot-based refill
// from a wave-local ray pool (the BVH2 kernel in trace.hip describes the
// scheme): a lane parks the first leaf it finds and keeps descending until
// the wave tests leaves together; the answer does not depend on the order.
// (Nearest hits stay on k_trace4_simple: on deep, heavily overlapping trees
// its 48-entry LDS stack beats a short LDS stack with global spill.)
// Src gives ray i: StoredRays reads it from HBM, AoRays makes AO sample i
// from its pixel's frame; both trace every ray in the launch's tree and
// write answer i at out[i].  ServeRays (kSegments) belongs to the server:
// ray i is ray `slot` of one request, which names its own tree and answer
// buffer.  A ray with an empty interval (tFar <= tNear, or a NaN bound) can
// pass no triangle's t test and finishes at once.
// A launch-wide ray source gives the ray index taken at queue position i
// (index) and that ray (ray).
struct StoredRays {
  static constexpr bool kSegments = false;
  const Ray* __restrict__ rays;
  __device__ uint64_t index(uint64_t i) const { return i; }
  __device__ Ray ray(uint64_t i) const { return rays[i]; }
};

struct AoRays {
  static constexpr bool kSegments = false;
  const AoFrame* __restrict__ frames;
  const uint16_t* __restrict__ order;  // rtao_order's direction order, or null
  uint32_t spp;
  float tmax;
  uint64_t seed;
  __device__ uint64_t index(uint64_t i) const {
    return order ? (i & ~uint64_t(kAoOrderRun - 1)) + order[i] : i;
  }
  // 32-bit division: the launcher keeps npix * spp below 2^32
  __device__ Ray ray(uint64_t i) const { return ao_ray(frames[uint32_t(i) / spp], i, tmax, seed); }
};

struct ServeRays {
  static constexpr bool kSegments = true;
  const uint32_t* __restrict__ off;  // nseg + 1 ray offsets
  const ServeSeg* __restrict__ segs;
  uint32_t nseg;
  __device__ Ray ray(uint64_t i, const void*& nodes, const TriPre*& tris, uint8_t*& out, uint64_t& slot) const {
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (off[mid] <= uint32_t(i)) lo = mid;
      else hi = mid;
    }
    const ServeSeg& S = segs[lo];
    nodes = S.nodes;
    tris = S.tris;
    out = static_cast<uint8_t*>(S.out);
    slot = i - off[lo];
    return S.rays[slot];
  }
};

template <bool kFp16, class Src, class Tri = MtTest>
__global__ __launch_bounds__(kBlock4, 8) void k_trace4_any_persistent(const void* __restrict__ quads,
                                                                   const TriPre* __restrict__ tris, Src src,
                                                                   uint64_t n, uint8_t* __restrict__ out,
                                                                   uint32_t* __restrict__ queue,
                                                                   uint32_t* __restrict__ spill) {
  __shared__ uint32_t stack_mem[kShortStack * kBlock4];
  const ShortStack<kShortStack, kBlock4> stk{stack_mem + threadIdx.x, spill + blockIdx.x * kBlock4 + threadIdx.x,
                                             gridDim.x * kBlock4};
  bool active = false;
  uint64_t rid = 0;
  float3 o = make_float3(0.f, 0.f, 0.f), d = o, inv = o, oi = o;
  bool fz = true;
  float tnear = 0.f;
  Tri tt;
  HitRec h;
  h.t = 0.f;
  h.prim = kInvalid;
  h.u = h.v = 0.f;
  int top = 0;
  uint32_t cur = kInvalid;
  uint32_t leaf = 0, leaf2 = 0;  // parked leaves (leaf2 only while leaf is held)
  RayPool pool;  // ray indices claimed by this wave (one queue: per-XCD queues measured 3 % slower here)
  // the tree and answer buffer of this lane's ray (a per-request source);
  // a launch-wide source keeps the kernel arguments
  const void* lane_quads = quads;
  const TriPre* lane_tris = tris;
  uint8_t* lane_out = out;

  while (true) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = __popcll(idle);
    if (!pool.drained() && (nidle >= kRefillAt4 || nidle == kWave)) {
      const uint64_t idx = pool.take<kClaim>(nidle, lanes_below(idle), queue, n);
      if (!active) {
        if (idx < n) {
          Ray r;
          if constexpr (Src::kSegments) {
            r = src.ray(idx, lane_quads, lane_tris, lane_out, rid);
          } else {
            rid = src.index(idx);
            r = src.ray(rid);
          }
          o = make_float3(r.pos_near.x, r.pos_near.y, r.pos_near.z);
          d = make_float3(r.dir_far.x, r.dir_far.y, r.dir_far.z);
          tnear = r.pos_near.w;
          h.t = fabsf(r.dir_far.w);
          h.prim = kInvalid;
          inv = safe_inverse(d);
          oi = neg_o_inv(o, inv);
          fz = fused_ok(oi);
          tt = Tri(o, d);
          top = 0;
          cur = (h.t > tnear) ? 0u : kInvalid;
          leaf = 0;
          leaf2 = 0;
          active = true;
        }
      }
    }
    if (__ballot(active) == 0) break;

    while (true) {
      // inner phase: park the first two leaves found; a lane with a parked
      // leaf keeps descending while enough lanes of the wave still look for
      // one
      while (true) {
        if (active && (cur & kLeafBit) != 0 && cur != kInvalid && (leaf == 0 || leaf2 == 0)) {
          if (leaf == 0) leaf = cur;
          else leaf2 = cur;
          cur = (top > 0) ? stk.pop(top) : kInvalid;
        }
        // leave for the leaf phase once at most kLeafPhaseAt lanes still
        // search and some lane holds a leaf (the slowest searchers would
        // otherwise keep the wave half idle)
        const bool searching = active && leaf == 0 && (cur & kLeafBit) == 0;
        const uint32_t nsearch = __popcll(__ballot(searching));
        if (nsearch == 0 || (nsearch <= kLeafPhaseAt && __any(leaf != 0))) break;
        if (active && (cur & kLeafBit) == 0) {
          // the children in stored order: the answer does not depend on the
          // order, and the distance sort costs a quarter of a visit's
          // instructions (on AO rays the sorted walk took 6 % longer, one
          // that only walks the nearest child first 3 % longer).  The first
          // entered child is walked next; the other entered ones are pushed,
          // the lowest on top.
          float2 bt[4];
          uint32_t ch[4];
          quad_boxes<kFp16, kFp16>(Src::kSegments ? lane_quads : quads, cur, o, inv, oi, fz, tt, bt, ch);
          bool hk[4];
#pragma unroll
          for (int k = 0; k < 4; k++) hk[k] = child_hit<!(kFp16 && !Tri::kRobustBoxes)>(bt[k], ch[k], tnear, h.t);
          stk.push_masked3(top, hk[3] && (hk[0] || hk[1] || hk[2]), ch[3], hk[2] && (hk[0] || hk[1]), ch[2],
                           hk[1] && hk[0], ch[1]);
          if (hk[0]) cur = ch[0];
          else if (hk[1]) cur = ch[1];
          else if (hk[2]) cur = ch[2];
          else if (hk[3]) cur = ch[3];
          else cur = (top > 0) ? stk.pop(top) : kInvalid;
        }
      }
      // leaf phase
      if (active && leaf != 0) {
        const TriPre* T = Src::kSegments ? lane_tris : tris;
        const uint32_t s = leaf_start(leaf), c = leaf_count(leaf);
        bool hit = false;
        for (uint32_t k = s; k < s + c && !hit; k++) hit = tri_hit_before(tt, o, d, tnear, T[k], h.t);
        if (!hit && leaf2 != 0) {
          const uint32_t s2 = leaf_start(leaf2), c2 = leaf_count(leaf2);
          for (uint32_t k = s2; k < s2 + c2 && !hit; k++) hit = tri_hit_before(tt, o, d, tnear, T[k], h.t);
        }
        if (hit) h.prim = 0u;  // any id but kInvalid: the query is answered
        leaf = 0;
        leaf2 = 0;
      }
      if (active && (h.prim != kInvalid || cur == kInvalid)) {
        (Src::kSegments ? lane_out : out)[rid] = (h.prim != kInvalid) ? 1 : 0;
        active = false;
      }
      const uint32_t nact = __popcll(__ballot(active));
      if (nact == 0 || (!pool.drained() && kWave - nact >= kRefillAt4)) break;
    }
  }
}

template <bool kFp16, class Src, class Tri>
hipError_t launch_any4(const Blas& b, Src src, uint64_t n, uint8_t* out, hipStream_t stream,
                       uint32_t* workspace = nullptr, size_t workspace_words = 0) {
  const void* kfn = reinterpret_cast<const void*>(&k_trace4_any_persistent<kFp16, Src, Tri>);
  const uint64_t want = (n + kBlock4 - 1) / kBlock4;
  constexpr size_t kLaneWords = kStackDepth4 - kShortStack;
  uint64_t cap = resident_grid(kfn, kBlock4);
  if (workspace) cap = std::min<uint64_t>(cap, (workspace_words - 64) / (kBlock4 * kLaneWords));
  const int grid = int(std::max<uint64_t>(1, std::min<uint64_t>(want, cap)));
  const void* q = kFp16 ? static_cast<const void*>(b.quads16) : static_cast<const void*>(b.quads32);
  // queue counter + spill area for stack entries beyond the LDS part
  const size_t spill_words = size_t(grid) * kBlock4 * kLaneWords;
  uint32_t* buf = workspace;
  hipError_t e = hipSuccess;
  if (!buf && (e = hipMallocAsync(reinterpret_cast<void**>(&buf), sizeof(uint32_t) * (64 + spill_words), stream)) !=
                  hipSuccess)
    return e;
  k_reset_counter4<<<1, 1, 0, stream>>>(buf);
  k_trace4_any_persistent<kFp16, Src, Tri><<<grid, kBlock4, 0, stream>>>(q, tri_records<Tri>(b), src, n, out, buf,
                                                                          buf + 64);
  e = hipGetLastError();
  hipError_t f = workspace ? hipSuccess : hipFreeAsync(buf, stream);
  return e != hipSuccess ? e : f;
}

// Sort key of a ray for coherent waves: 7 bits of direction (the cube face of
// the largest |d| component, 6 faces, times a 4x4 grid of the other two
// components divided by it) above a 24-bit Morton code of the origin on a
// 256^3 grid over the scene box (clamped).  Rays with similar directions and
// nearby origins then share waves, so their node and triangle fetches hit the
// same cache lines.  A NaN component lands in cell 0.
__device__ inline uint32_t spread8(uint32_t x) {
  x = (x | (x << 8)) & 0x0000F00Fu;
  x = (x | (x << 4)) & 0x000C30C3u;
  x = (x | (x << 2)) & 0x00249249u;
  return x;
}

__global__ __launch_bounds__(256) void k_ray_keys(const Ray* __restrict__ rays, uint32_t n, float3 lo, float3 scale,
                                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ ids) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const Ray r = rays[i];
  const float dx = r.dir_far.x, dy = r.dir_far.y, dz = r.dir_far.z;
  const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
  uint32_t face;
  float m, p, q;
  if (ax >= ay && ax >= az) {
    face = dx < 0.0f ? 1u : 0u;
    m = ax, p = dy, q = dz;
  } else if (ay >= az) {
    face = dy < 0.0f ? 3u : 2u;
    m = ay, p = dx, q = dz;
  } else {
    face = dz < 0.0f ? 5u : 4u;
    m = az, p = dx, q = dy;
  }
  const float inv_m = m > 0.0f ? 1.0f / m : 0.0f;
  const uint32_t cp = uint32_t(fminf(fmaxf((p * inv_m + 1.0f) * 2.0f, 0.0f), 3.0f));
  const uint32_t cq = uint32_t(fminf(fmaxf((q * inv_m + 1.0f) * 2.0f, 0.0f), 3.0f));
  const uint32_t cx = uint32_t(fminf(fmaxf((r.pos_near.x - lo.x) * scale.x, 0.0f), 255.0f));
  const uint32_t cy = uint32_t(fminf(fmaxf((r.pos_near.y - lo.y) * scale.y, 0.0f), 255.0f));
  const uint32_t cz = uint32_t(fminf(fmaxf((r.pos_near.z - lo.z) * scale.z, 0.0f), 255.0f));
  keys[i] = ((face * 16u + cp * 4u + cq) << 24) | spread8(cx) | (spread8(cy) << 1) | (spread8(cz) << 2);
  ids[i] = i;
}

// The order the trace kernel reads: the sorted one, or none (unsorted) when
// the sort failed (its look-back gave up: *err != 0) and left it incomplete.
__global__ void k_order_or_null(const uint32_t* __restrict__ err, const uint32_t* order, const uint32_t** slot) {
  *slot = (*err == 0u) ? order : nullptr;
}

template <bool kFp16, class Tri>
hipError_t launch_simple4(const Blas& b, const TraceArgs& a, const uint32_t* const* order_slot,
                          hipStream_t stream) {
  const uint32_t blocks = uint32_t((a.n + kBlock4 - 1) / kBlock4);
  const void* q = kFp16 ? static_cast<const void*>(b.quads16) : static_cast<const void*>(b.quads32);
  // stack entries beyond the LDS part, per lane (none while LDS holds them all)
  uint32_t* spill = nullptr;
  hipError_t e = hipSuccess;
  if (kSimpleLds4 < kStackDepth4 &&
      (e = hipMallocAsync(reinterpret_cast<void**>(&spill),
                          sizeof(uint32_t) * size_t(blocks) * kBlock4 * (kStackDepth4 - kSimpleLds4), stream)) !=
          hipSuccess)
    return e;
  const uint2 ids = make_uint2(a.inst_id, a.geom_id);
  const TriPre* tris = tri_records<Tri>(b);
  if (a.anyhit) {
    if (a.stats)
      k_trace4_simple<true, kFp16, true, Tri><<<blocks, kBlock4, 0, stream>>>(q, tris, a.rays, a.n, a.out, a.format,
                                                                             ids, a.stats, spill, nullptr);
    else
      k_trace4_simple<true, kFp16, false, Tri><<<blocks, kBlock4, 0, stream>>>(q, tris, a.rays, a.n, a.out, a.format,
                                                                              ids, nullptr, spill, nullptr);
  } else {
    if (a.stats)
      k_trace4_simple<false, kFp16, true, Tri><<<blocks, kBlock4, 0, stream>>>(q, tris, a.rays, a.n, a.out, a.format,
                                                                              ids, a.stats, spill, nullptr);
    else
      k_trace4_simple<false, kFp16, false, Tri><<<blocks, kBlock4, 0, stream>>>(q, tris, a.rays, a.n, a.out,
                                                                               a.format, ids, nullptr, spill, order_slot);
  }
  e = hipGetLastError();
  const hipError_t f = spill ? hipFreeAsync(spill, stream) : hipSuccess;
  return e != hipSuccess ? e : f;
}

// Nearest hits of a large batch: the rays are sorted into coherent waves
// first (k_ray_keys, then the onesweep radix sort of the engine's builder),
// and each thread traces the ray the order gives it; the answers land at the
// rays' own indices, so the output is what an unsorted trace writes.
template <bool kFp16, class Tri>
hipError_t launch_nearest4_sorted(const Blas& b, const TraceArgs& a, hipStream_t stream) {
  const uint64_t n = a.n;
  const size_t sort_words = radix_sort_scratch_words(n, 4);
  // keys, ids (twice), radix scratch, the error word and the order slot
  const size_t words = 4 * n + sort_words + 64;
  uint32_t* base = nullptr;
  hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&base), sizeof(uint32_t) * words, stream);
  if (e != hipSuccess) return e;
  // base[0..1]: the order slot (8-byte aligned), base[2]: the sort's error word
  const uint32_t** slot = reinterpret_cast<const uint32_t**>(base);
  uint32_t* err = base + 2;
  uint32_t *ka = base + 64, *va = ka + n, *kb = ka + 2 * n, *vb = ka + 3 * n;
  uint32_t* scratch = ka + 4 * n;
  const float3 lo = make_float3(b.box[0], b.box[1], b.box[2]);
  const float ex = b.box[4] - b.box[0], ey = b.box[5] - b.box[1], ez = b.box[6] - b.box[2];
  const float3 scale = make_float3(ex > 0.f ? 256.f / ex : 0.f, ey > 0.f ? 256.f / ey : 0.f, ez > 0.f ? 256.f / ez : 0.f);
  e = fill_words(err, 1, 0u, stream);
  if (e == hipSuccess) {
    k_ray_keys<<<uint32_t((n + 255) / 256), 256, 0, stream>>>(a.rays, uint32_t(n), lo, scale, ka, va);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = radix_sort(ka, va, kb, vb, n, 0, 4, scratch, err, stream);
  const uint32_t* order = va;  // four passes: the sorted pairs end in (ka, va)
  if (e == hipSuccess) {
    k_order_or_null<<<1, 1, 0, stream>>>(err, order, slot);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = launch_simple4<kFp16, Tri>(b, a, slot, stream);
  const hipError_t f = hipFreeAsync(base, stream);
  return e != hipSuccess ? e : f;
}

template <class Tri>
hipError_t trace_blas4_with(const Blas& b, const TraceArgs& a, bool fp16, hipStream_t stream) {
  if (a.anyhit && !a.stats) {
    const StoredRays src{a.rays};
    uint8_t* out = static_cast<uint8_t*>(a.out);
    return fp16 ? launch_any4<true, StoredRays, Tri>(b, src, a.n, out, stream, a.workspace, a.workspace_words)
                : launch_any4<false, StoredRays, Tri>(b, src, a.n, out, stream, a.workspace, a.workspace_words);
  }
  if (!a.stats && !a.anyhit && a.n >= kSortRaysMin4)
    return fp16 ? launch_nearest4_sorted<true, Tri>(b, a, stream) : launch_nearest4_sorted<false, Tri>(b, a, stream);
  return fp16 ? launch_simple4<true, Tri>(b, a, nullptr, stream) : launch_simple4<false, Tri>(b, a, nullptr, stream);
}

}  // namespace

hipError_t trace_ao4(const Blas& b, bool fp16, bool watertight, const AoFrame* frames, const uint16_t* order,
                     uint64_t npix, uint32_t spp, float tmax, uint64_t seed, uint8_t* out, hipStream_t stream) {
  const uint64_t n = npix * spp;
  if (n == 0) return hipSuccess;
  if (!b.quads32 || !b.quads16 || n > 0xFFFFFFFFull || (watertight && !b.tris_wt)) return hipErrorInvalidValue;
  const AoRays src{frames, order, spp, tmax, seed};
  if (watertight)
    return fp16 ? launch_any4<true, AoRays, WtTest>(b, src, n, out, stream)
                : launch_any4<false, AoRays, WtTest>(b, src, n, out, stream);
  return fp16 ? launch_any4<true, AoRays, MtTest>(b, src, n, out, stream)
              : launch_any4<false, AoRays, MtTest>(b, src, n, out, stream);
}

hipError_t serve_any4(bool fp16, const uint32_t* off, const ServeSeg* segs, uint32_t nseg, uint32_t nrays,
                      uint32_t grid, uint32_t* queue, uint32_t* spill, hipStream_t stream) {
  if (nrays == 0) return hipSuccess;
  const ServeRays src{off, segs, nseg};
  grid = std::max<uint32_t>(1, std::min<uint32_t>(grid, (nrays + kBlock4 - 1) / kBlock4));
  if (fp16)
    k_trace4_any_persistent<true, ServeRays, MtTest><<<grid, kBlock4, 0, stream>>>(nullptr, nullptr, src, nrays,
                                                                                   nullptr, queue, spill);
  else
    k_trace4_any_persistent<false, ServeRays, MtTest><<<grid, kBlock4, 0, stream>>>(nullptr, nullptr, src, nrays,
                                                                                    nullptr, queue, spill);
  return hipGetLastError();
}

void warm_any_hit_launchers() {
  warm_bvh2_any_hit_launchers();
  const void* k[] = {reinterpret_cast<const void*>(&k_trace4_any_persistent<true, StoredRays, MtTest>),
                     reinterpret_cast<const void*>(&k_trace4_any_persistent<false, StoredRays, MtTest>),
                     reinterpret_cast<const void*>(&k_trace4_any_persistent<true, StoredRays, WtTest>),
                     reinterpret_cast<const void*>(&k_trace4_any_persistent<false, StoredRays, WtTest>)};
  for (const void* f : k) (void)resident_grid(f, kBlock4);
}

uint32_t serve_any4_resident_grid(bool fp16) {
  return fp16 ? resident_grid(reinterpret_cast<const void*>(&k_trace4_any_persistent<true, ServeRays, MtTest>), kBlock4)
              : resident_grid(reinterpret_cast<const void*>(&k_trace4_any_persistent<false, ServeRays, MtTest>),
                              kBlock4);
}

hipError_t trace_blas4(const Blas& b, const TraceArgs& a, bool fp16, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  if (!b.quads32 || !b.quads16) return hipErrorInvalidValue;
  if (a.watertight) {
    if (!b.tris_wt) return hipErrorInvalidValue;
    return trace_blas4_with<WtTest>(b, a, fp16, stream);
  }
  return trace_blas4_with<MtTest>(b, a, fp16, stream);
}

}  // namespace hrt
