// hrt_device.h -- device-side formats and arithmetic shared by the builder and
// the traversal kernels (gfx950, wave64).
//
// Every floating-point expression here mirrors the operation order of the
// reference so that results are bit-identical to the CPU path when the file
// is compiled with -ffp-contract=off (no fused multiply-add) and with the
// default correctly rounded fp32 division:
//   slab test           core/raytrace_common.h:67-80
//   safe reciprocal     core/raytrace_common.h:57-65
//   Moller-Trumbore     core/BVH2CommonRT.cpp:26-37 (1e-6 slack, :37)
//   leaf packing        core/raytrace_common.h:38-55
//
// min/max use v_min_f32/v_max_f32.  For finite rays and boxes they return
// exactly what std::min/std::max return, except possibly the sign of a zero;
// the slab results are only ever compared (<=, <, >=), where +0 == -0, so the
// traversal decisions and the hits are unaffected.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hrt {

constexpr uint32_t kStartMask = 0x00FFFFFFu;
constexpr uint32_t kSizeMask = 0x7F000000u;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr uint32_t kEmptyLeaf = 0x80000000u;  // PackOffsetAndSize(0, 0): a leaf without triangles
// Deepest possible BLAS path: the Karras tree over unique 30-bit codes has at
// most 30 internal levels, so at most 30 deferred children are ever pending.
constexpr int kStackDepth = 32;

__host__ __device__ inline uint32_t pack_leaf(uint32_t start, uint32_t count) {
  return kLeafBit | ((count << 24) & kSizeMask) | (start & kStartMask);
}
__host__ __device__ inline uint32_t leaf_start(uint32_t v) { return v & kStartMask; }
__host__ __device__ inline uint32_t leaf_count(uint32_t v) { return (v & kSizeMask) >> 24; }

// ---------------------------------------------------------------------------
// HBM layouts
// ---------------------------------------------------------------------------
// BVH2 node pair: the two children of one internal Karras node, 64 bytes,
// fetched together on every visit (four 16-byte loads).  Pair k holds the
// children of internal node k, so the pair array is indexed by the Karras
// internal-node id and is laid out in Morton order.  A child reference is
// either an internal node id (pair index) or a packed leaf (kLeafBit set).
struct __attribute__((aligned(64))) NodePair {
  float4 lmin;  // xyz = left child box min,  w = bits of left child ref
  float4 lmax;  // xyz = left child box max,  w = unused (0)
  float4 rmin;  // xyz = right child box min, w = bits of right child ref
  float4 rmax;  // xyz = right child box max, w = unused (0)
};

// Triangle in leaf order, 48 bytes: vertex A, the edges B-A and C-A as the
// reference computes them at test time (BVH2CommonRT.cpp:26-27; the
// subtraction is exact-once either way, so storing it is bit-identical), and
// the original triangle id in a.w.
struct __attribute__((aligned(16))) TriPre {
  float4 a;   // xyz = A, w = bits(original triangle id)
  float4 e1;  // xyz = B - A
  float4 e2;  // xyz = C - A
};

// Per-lane traversal stack of a persistent kernel: the first kLds entries in
// LDS (entry k of the lane at lds[k * kStride]), deeper entries in a global
// spill area that the lane owns for the whole launch (entry k at
// spill[(k - kLds) * spill_stride]).  Typical traversals stay well inside the
// LDS part; the spill area makes the proven worst-case depth safe while the
// LDS footprint stays small enough for full occupancy.
// The LDS part is held as an LDS-address-space pointer: a pop then compiles
// to a ds_read or a global load on its own branch, not to one flat load
// through a pointer chosen between the two.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
template <int kLds, int kStride>
struct ShortStack {
  lds_u32* lds;
  uint32_t* spill;
  uint32_t spill_stride;
  __device__ ShortStack(uint32_t* lds_generic, uint32_t* spill_, uint32_t stride)
      : lds((lds_u32*)(lds_generic)), spill(spill_), spill_stride(stride) {}
  __device__ inline void push(int& top, uint32_t v) const {
    if (top < kLds) lds[top * kStride] = v;
    else spill[size_t(top - kLds) * spill_stride] = v;
    top++;
  }
  __device__ inline uint32_t pop(int& top) const {
    --top;
    if (top < kLds) return lds[top * kStride];
    return spill[size_t(top - kLds) * spill_stride];
  }
  // Pushes c3 if p3, then c2 if p2, then c1 if p1.  With room for three more
  // entries in the LDS part, each value is stored at the top slot whether or
  // not it is pushed (a slot at or above top holds nothing), so there is no
  // branch.
  __device__ inline void push_masked3(int& top, bool p3, uint32_t c3, bool p2, uint32_t c2, bool p1,
                                      uint32_t c1) const {
    if (top + 3 <= kLds) {
      lds[top * kStride] = c3;
      top += int(p3);
      lds[top * kStride] = c2;
      top += int(p2);
      lds[top * kStride] = c1;
      top += int(p1);
    } else {
      if (p3) push(top, c3);
      if (p2) push(top, c2);
      if (p1) push(top, c1);
    }
  }
};
constexpr int kShortStack = 16;  // LDS entries per lane in the persistent kernels

// The logical block a workgroup works on, so that each XCD's L2 serves one
// contiguous stretch of the rays.  The dispatcher hands workgroups to the
// eight XCDs in turn (workgroup b to XCD b mod 8); XCD x gets the x-th
// eighth of the logical blocks, in order.  A permutation of [0, n).
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t n) {
  const uint32_t per = n / kXcds, rem = n % kXcds;
  const uint32_t x = b % kXcds, i = b / kXcds;
  return x * per + (x < rem ? x : rem) + i;
}

// Set bits of m in the lanes below this one (v_mbcnt_lo / v_mbcnt_hi).
__device__ inline uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

// Ray indices for the idle lanes of a persistent wave.  The wave claims
// kClaim consecutive indices from the global queue with one atomic and hands
// them out to its own idle lanes at later refills, so the single global
// counter is hit once per kClaim rays instead of once per refill (a counter
// that every wave of the chip hits at each refill limits the whole kernel).
// Every member is wave-uniform; take() must be called by all lanes of the
// wave together.
struct RayPool {
  uint64_t next = 0, end = 0;  // indices claimed by the wave, not yet given out
  bool exhausted = false;      // the global queue has been run past n

  // Returns this lane's ray index, or n when it gets none.  nidle = number of
  // idle lanes, rank = this lane's position among them.
  template <uint32_t kClaim>
  __device__ inline uint64_t take(uint32_t nidle, uint32_t rank, uint32_t* queue, uint64_t n) {
    const uint64_t avail = end - next;
    const bool claim = avail < nidle && !exhausted;
    uint64_t nb = 0;
    if (claim) {
      uint32_t b = 0;
      if ((threadIdx.x & 63) == 0) b = atomicAdd(queue, kClaim);
      nb = uint64_t(__shfl(b, 0));
    }
    uint64_t idx = n;
    if (rank < avail) idx = next + rank;
    else if (claim) idx = nb + (rank - avail);
    if (claim) {
      end = (nb + kClaim < n) ? nb + kClaim : n;
      next = nb + (nidle - avail);
      if (next > end) next = end;
      if (nb + kClaim >= n) exhausted = true;
    } else {
      next += (avail < nidle) ? avail : nidle;
    }
    return idx < n ? idx : n;
  }
  __device__ inline bool drained() const { return exhausted && next >= end; }
};

// RayPool with one queue per XCD: [0, n) is cut into kXcds contiguous
// eighths, and a wave claims from the eighth of its own XCD (workgroup b runs
// on XCD b mod kXcds) until that is used up, then from the next eighths in
// turn, so each XCD's L2 mostly serves one stretch of the rays.  The queue
// holds the eighths' counters kXcdQueueStride words apart (all zero before
// the launch).  Every member is wave-uniform; take() must be called by all
// lanes of the wave together.
constexpr uint32_t kXcdQueueStride = 8;
constexpr uint32_t kXcdQueueWords = kXcds * kXcdQueueStride;
struct XcdRayPool {
  uint64_t next = 0, end = 0;  // indices claimed by the wave, not yet given out
  uint32_t part = 0;           // the eighth being claimed from
  uint32_t tried = 0;          // eighths found used up, from the wave's own one on
  bool exhausted = false;      // every eighth is used up

  __device__ explicit XcdRayPool(uint32_t block) : part(block % kXcds) {}

  template <uint32_t kClaim>
  __device__ inline uint64_t take(uint32_t nidle, uint32_t rank, uint32_t* queue, uint64_t n) {
    const uint64_t avail = end - next;
    uint64_t idx = (rank < avail) ? next + rank : n;
    if (avail >= nidle || exhausted) {
      next += (avail < nidle) ? avail : nidle;
      return idx < n ? idx : n;
    }
    next = end;  // the first avail lanes took the rest
    while (tried < kXcds) {
      const uint32_t p = (part + tried) % kXcds;
      const uint64_t lo = n * p / kXcds, hi = n * (p + 1) / kXcds;
      uint32_t b = 0;
      if ((threadIdx.x & 63) == 0) b = atomicAdd(queue + p * kXcdQueueStride, kClaim);
      const uint64_t nb = lo + uint64_t(__shfl(b, 0));
      if (nb < hi) {
        end = (nb + kClaim < hi) ? nb + kClaim : hi;
        if (rank >= avail && nb + (rank - avail) < end) idx = nb + (rank - avail);
        next = nb + (nidle - avail);
        if (next > end) next = end;
        return idx < n ? idx : n;
      }
      tried++;
    }
    exhausted = true;
    return idx < n ? idx : n;
  }
  __device__ inline bool drained() const { return exhausted && next >= end; }
};

// BVH4 node with FP32 boxes, 128 bytes, coordinates of the four children
// grouped per axis so each 16-byte load feeds four slab tests.  Children are
// the grandchildren of one even-depth Karras node in left-to-right order;
// an unused slot has the inverted box (+1e9, -1e9) and the empty leaf
// 0x80000000 (cbvh_core.cpp:402-409).  An internal child reference is a quad
// index.
struct __attribute__((aligned(128))) Quad32 {
  float4 lox, hix, loy, hiy, loz, hiz;
  uint4 ref;
  uint4 pad;
};

// BVH4 node with FP16 boxes, 64 bytes: per child one uint4 as the reference's
// BVH4HalfRT stores it (half2 {minx, miny}, {minz, maxx}, {maxy, maxz}, ref;
// cbvh_fat.cpp:477-497), minima rounded to nearest then one ulp down, maxima
// one ulp up (cbvh_fat.cpp:447-448).
struct __attribute__((aligned(64))) Quad16 {
  uint4 c[4];
};

// Deepest BVH4 path: at most kMaxLevels4 / 2 node levels (the even depths of
// a BVH2 of at most kMaxLevels4 internal levels), at most 3 deferred children
// per level.  An LBVH over unique 30-bit codes has at most 30 internal
// levels; the clustered second tree is kept only within kMaxLevels4 (its
// build measures the depth: 38 levels on a 4 M-triangle random soup).
constexpr int kStackDepth4 = 63;
constexpr int kMaxLevels4 = 2 * (kStackDepth4 / 3);

struct Ray {  // 32-byte ray record: posAndNear, dirAndFar (CrossRT.h:143)
  float4 pos_near;
  float4 dir_far;
};

// ---------------------------------------------------------------------------
// Arithmetic
// ---------------------------------------------------------------------------
__device__ inline float3 safe_inverse(float3 d) {
  const float eps = 1.0e-36f;
  float3 r;
  r.x = 1.0f / (fabsf(d.x) > eps ? d.x : copysignf(eps, d.x));
  r.y = 1.0f / (fabsf(d.y) > eps ? d.y : copysignf(eps, d.y));
  r.z = 1.0f / (fabsf(d.z) > eps ? d.z : copysignf(eps, d.z));
  return r;
}

__device__ inline float2 slab(float3 o, float3 inv, float bminx, float bminy, float bminz, float bmaxx,
                              float bmaxy, float bmaxz) {
  const float lo = inv.x * (bminx - o.x);
  const float hi = inv.x * (bmaxx - o.x);
  const float lo1 = inv.y * (bminy - o.y);
  const float hi1 = inv.y * (bmaxy - o.y);
  const float lo2 = inv.z * (bminz - o.z);
  const float hi2 = inv.z * (bmaxz - o.z);
  const float tmin = fmaxf(fminf(lo, hi), fminf(lo1, hi1));
  const float tmax = fminf(fmaxf(lo, hi), fmaxf(lo1, hi1));
  return make_float2(fmaxf(tmin, fminf(lo2, hi2)), fminf(tmax, fmaxf(lo2, hi2)));
}

// The box test of the watertight mode, made conservative as Ize shows
// ("Robust BVH Ray Traversal", JCGT 2(2), 2013): a plane distance
// inv * (b - o) carries three roundings (the reciprocal, the difference, the
// product), a relative error within gamma3 = 3u / (1 - 3u), u = 2^-24, so the
// exit distance is scaled by 1 + 2 gamma3, taken as the next float above it
// (1 + 4 * 2^-23), and a box the exact ray meets is never rejected.  An axis
// whose direction component lies within +-1e-36 (where safe_inverse puts
// +-1e36 in place of 1/d) is parallel: its slab is every t when
// lo <= o <= hi and no t otherwise, so a ray running inside a box face keeps
// the box.
constexpr float kRobustExit = 1.000000476837158203125f;
__device__ inline uint32_t parallel_axes(float3 d) {
  return (fabsf(d.x) > 1.0e-36f ? 0u : 1u) | (fabsf(d.y) > 1.0e-36f ? 0u : 2u) | (fabsf(d.z) > 1.0e-36f ? 0u : 4u);
}
__device__ inline void robust_axis(float o, float inv, bool par, float lo, float hi, float& a, float& b) {
  const float t0 = inv * (lo - o), t1 = inv * (hi - o);
  a = fminf(t0, t1);
  b = fmaxf(t0, t1);
  if (par) {
    const bool in = lo <= o && o <= hi;
    a = in ? -__builtin_inff() : __builtin_inff();
    b = in ? __builtin_inff() : -__builtin_inff();
  }
}
__device__ inline float2 slab_robust(float3 o, float3 inv, uint32_t par, float bminx, float bminy, float bminz,
                                     float bmaxx, float bmaxy, float bmaxz) {
  float ax, bx, ay, by, az, bz;
  robust_axis(o.x, inv.x, (par & 1u) != 0u, bminx, bmaxx, ax, bx);
  robust_axis(o.y, inv.y, (par & 2u) != 0u, bminy, bmaxy, ay, by);
  robust_axis(o.z, inv.z, (par & 4u) != 0u, bminz, bmaxz, az, bz);
  return make_float2(fmaxf(fmaxf(ax, ay), az), fminf(fminf(bx, by), bz) * kRobustExit);
}

// -(o * inv) per axis, made once per ray for slab_fused.
__device__ inline float3 neg_o_inv(float3 o, float3 inv) {
  return make_float3(-(o.x * inv.x), -(o.y * inv.y), -(o.z * inv.z));
}

// slab_fused needs a finite -(o * inv) on every axis.  It overflows for an
// axis-parallel direction (safe_inverse gives +-1e36) with |o| above ~340 on
// that axis; such a ray keeps the unfused slab.
__device__ inline bool fused_ok(float3 oi) {
  return __builtin_isfinite(oi.x) && __builtin_isfinite(oi.y) && __builtin_isfinite(oi.z);
}

struct HitRec {
  float t;
  uint32_t prim;  // original triangle id, kInvalid on miss
  float u, v;
};

// Moller-Trumbore arithmetic in the operation order of BVH2CommonRT.cpp:26-37:
// leaves t, u, v and says whether the barycentric slack (1e-6) and the
// t > tNear condition hold.  Callers add their own t-acceptance rule.
__device__ inline bool mt_eval(float3 o, float3 d, float tnear, const TriPre& tri, float& t, float& u, float& v) {
  const float3 A = make_float3(tri.a.x, tri.a.y, tri.a.z);
  const float3 e1 = make_float3(tri.e1.x, tri.e1.y, tri.e1.z);
  const float3 e2 = make_float3(tri.e2.x, tri.e2.y, tri.e2.z);
  // pvec = cross(dir, e2)
  const float px = d.y * e2.z - d.z * e2.y;
  const float py = d.z * e2.x - d.x * e2.z;
  const float pz = d.x * e2.y - d.y * e2.x;
  // tvec = pos - A
  const float tx = o.x - A.x;
  const float ty = o.y - A.y;
  const float tz = o.z - A.z;
  // qvec = cross(tvec, e1)
  const float qx = ty * e1.z - tz * e1.y;
  const float qy = tz * e1.x - tx * e1.z;
  const float qz = tx * e1.y - ty * e1.x;
  const float invDet = 1.0f / (e1.x * px + e1.y * py + e1.z * pz);
  v = (tx * px + ty * py + tz * pz) * invDet;
  u = (qx * d.x + qy * d.y + qz * d.z) * invDet;
  t = (e2.x * qx + e2.y * qy + e2.z * qz) * invDet;
  return v >= -1e-6f && u >= -1e-6f && (u + v <= 1.0f + 1e-6f) && t > tnear;
}

// The reference's rule: the first strictly closer hit wins (BVH2CommonRT.cpp:37).
__device__ inline bool mt_test(float3 o, float3 d, float tnear, const TriPre& tri, HitRec& h) {
  float t, u, v;
  if (mt_eval(o, d, tnear, tri, t, u, v) && t < h.t) {
    h.t = t;
    h.prim = __float_as_uint(tri.a.w);
    h.u = u;
    h.v = v;
    return true;
  }
  return false;
}

// Same test with the (t, primId) rule of the BVH4 path: an equal-distance
// triangle wins when its original id is smaller, so the nearest hit does not
// depend on the order nodes are visited in.  t == tFar with no hit yet stays
// outside the open interval.
__device__ inline bool mt_test_lex(float3 o, float3 d, float tnear, const TriPre& tri, HitRec& h) {
  float t, u, v;
  const uint32_t prim = __float_as_uint(tri.a.w);
  if (mt_eval(o, d, tnear, tri, t, u, v) &&
      (t < h.t || (t == h.t && h.prim != kInvalid && prim < h.prim))) {
    h.t = t;
    h.prim = prim;
    h.u = u;
    h.v = v;
    return true;
  }
  return false;
}

// ---------------------------------------------------------------------------
// Triangle-test policies.  The traversal kernels take one as a template
// parameter; a lane makes its ray's policy object once (per refill) and tests
// every triangle of its leaves with it.
//   MtTest  the reference's Moller-Trumbore with the 1e-6 barycentric slack
//           (mt_eval) on TriPre records {A, id}, B-A, C-A.
//   WtTest  the watertight test of Woop, Benthin and Wald ("Watertight
//           Ray/Triangle Intersection", JCGT 2(1), 2013) in FP32 on records
//           {A, id}, B, C (absolute vertices).  The ray's largest direction
//           axis becomes z (lowest axis on ties; x and y swapped when that
//           component is negative, which keeps the winding), the vertices
//           are moved to the ray origin and sheared so the ray runs along z,
//           and the three 2D edge functions U, V, W decide: all >= 0 or all
//           <= 0 is a hit.  Two triangles that share an edge evaluate its
//           function from the same sheared vertices with the operands
//           exchanged, so they get exactly opposite values (or both zero):
//           no ray passes between them, with no slack.  An edge value of
//           zero counts as inside on both sides (the paper recomputes zeros
//           in double; this FP32 form leaves such a ray to both triangles,
//           which keeps the surface closed).  det = U + V + W; a zero det
//           (a degenerate triangle, or the ray in its plane) is a miss;
//           t = T / det with T = U Az + V Bz + W Cz, coords[0] = W / det
//           (the weight of C) and coords[1] = V / det (the weight of B), the
//           same meaning as the reference's u and v.
// ---------------------------------------------------------------------------
// A triangle-test policy also carries its box test: the reference's slab for
// its Moller-Trumbore, the conservative slab_robust for the watertight test
// (a watertight triangle test behind a box test that rounds a touched box
// away would still let rays through).
struct MtTest {
  static constexpr bool kRobustBoxes = false;
  __device__ MtTest() {}
  __device__ MtTest(float3, float3) {}
  __device__ bool eval(float3 o, float3 d, float tnear, const TriPre& tri, float& t, float& u, float& v) const {
    return mt_eval(o, d, tnear, tri, t, u, v);
  }
  __device__ float2 box(float3 o, float3 inv, float bminx, float bminy, float bminz, float bmaxx, float bmaxy,
                        float bmaxz) const {
    return slab(o, inv, bminx, bminy, bminz, bmaxx, bmaxy, bmaxz);
  }
};

__device__ inline float axis_of(float3 v, uint32_t k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

struct WtTest {
  static constexpr bool kRobustBoxes = true;
  uint32_t axes;  // kx | ky << 2 | kz << 4 | parallel_axes(d) << 6
  float sx, sy, sz;
  __device__ WtTest() : axes(0), sx(0.f), sy(0.f), sz(0.f) {}
  __device__ WtTest(float3, float3 d) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const uint32_t kz = (ax >= ay) ? (ax >= az ? 0u : 2u) : (ay >= az ? 1u : 2u);
    uint32_t kx = kz == 2u ? 0u : kz + 1u;
    uint32_t ky = kx == 2u ? 0u : kx + 1u;
    const float dz = axis_of(d, kz);
    if (dz < 0.0f) {
      const uint32_t k = kx;
      kx = ky;
      ky = k;
    }
    axes = kx | (ky << 2) | (kz << 4) | (parallel_axes(d) << 6);
    sx = axis_of(d, kx) / dz;
    sy = axis_of(d, ky) / dz;
    sz = 1.0f / dz;
  }
  __device__ bool eval(float3 o, float3, float tnear, const TriPre& tri, float& t, float& u, float& v) const {
    const uint32_t kx = axes & 3u, ky = (axes >> 2) & 3u, kz = (axes >> 4) & 3u;
    const float3 A = make_float3(tri.a.x - o.x, tri.a.y - o.y, tri.a.z - o.z);
    const float3 B = make_float3(tri.e1.x - o.x, tri.e1.y - o.y, tri.e1.z - o.z);
    const float3 C = make_float3(tri.e2.x - o.x, tri.e2.y - o.y, tri.e2.z - o.z);
    const float Akz = axis_of(A, kz), Bkz = axis_of(B, kz), Ckz = axis_of(C, kz);
    const float Ax = axis_of(A, kx) - sx * Akz, Ay = axis_of(A, ky) - sy * Akz;
    const float Bx = axis_of(B, kx) - sx * Bkz, By = axis_of(B, ky) - sy * Bkz;
    const float Cx = axis_of(C, kx) - sx * Ckz, Cy = axis_of(C, ky) - sy * Ckz;
    const float U = Cx * By - Cy * Bx;
    const float V = Ax * Cy - Ay * Cx;
    const float W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T = U * (sz * Akz) + V * (sz * Bkz) + W * (sz * Ckz);
    const float rcp = 1.0f / det;
    t = T * rcp;
    u = W * rcp;
    v = V * rcp;
    return t > tnear;
  }
  __device__ float2 box(float3 o, float3 inv, float bminx, float bminy, float bminz, float bmaxx, float bmaxy,
                        float bmaxz) const {
    return slab_robust(o, inv, axes >> 6, bminx, bminy, bminz, bmaxx, bmaxy, bmaxz);
  }
};

// The two acceptance rules with a triangle-test policy: the first strictly
// closer hit (BVH2), and (t, primId) (BVH4).
template <class Tri>
__device__ inline bool tri_test(const Tri& k, float3 o, float3 d, float tnear, const TriPre& tri, HitRec& h) {
  float t, u, v;
  if (k.eval(o, d, tnear, tri, t, u, v) && t < h.t) {
    h.t = t;
    h.prim = __float_as_uint(tri.a.w);
    h.u = u;
    h.v = v;
    return true;
  }
  return false;
}

// An any-hit query's test: whether the triangle is hit in (tnear, tfar).  The
// query ends at its first hit, so this equals tri_test / tri_test_lex with
// the incoming distance tfar, without keeping t, u, v or the id.
template <class Tri>
__device__ inline bool tri_hit_before(const Tri& k, float3 o, float3 d, float tnear, const TriPre& tri, float tfar) {
  float t, u, v;
  return k.eval(o, d, tnear, tri, t, u, v) && t < tfar;
}

template <class Tri>
__device__ inline bool tri_test_lex(const Tri& k, float3 o, float3 d, float tnear, const TriPre& tri, HitRec& h) {
  float t, u, v;
  const uint32_t prim = __float_as_uint(tri.a.w);
  if (k.eval(o, d, tnear, tri, t, u, v) && (t < h.t || (t == h.t && h.prim != kInvalid && prim < h.prim))) {
    h.t = t;
    h.prim = prim;
    h.u = u;
    h.v = v;
    return true;
  }
  return false;
}

// Two-level rule: inside one instance the first strictly closer hit wins, as
// in mt_test; a hit at the same distance in an instance with a smaller id
// replaces one from a larger id, so across instances the answer does not
// depend on the order the top-level tree is walked in.
__device__ inline bool mt_test_inst(float3 o, float3 d, float tnear, const TriPre& tri, uint32_t inst, HitRec& h,
                                    uint32_t& h_inst) {
  float t, u, v;
  if (mt_eval(o, d, tnear, tri, t, u, v) &&
      (t < h.t || (t == h.t && h.prim != kInvalid && inst < h_inst))) {
    h.t = t;
    h.prim = __float_as_uint(tri.a.w);
    h.u = u;
    h.v = v;
    h_inst = inst;
    return true;
  }
  return false;
}

// One instance as the two-level traversal reads it, 80 bytes.  inv_row[r] is
// row r of the inverse (world -> object) matrix; a point p maps to
//   p'.r = ((inv_row[r].x * p.x + inv_row[r].y * p.y) + inv_row[r].z * p.z) + inv_row[r].w * w
// with w = 1 for points and w = 0 for directions (matmul4x3 / matmul3x3,
// raytrace_common.h:82-90; the direction is not renormalised, so t keeps its
// world-space meaning).
struct __attribute__((aligned(16))) InstDev {
  float4 inv_row[3];
  const NodePair* pairs;
  const TriPre* tris;
  uint32_t geom;
  uint32_t inst;
  uint32_t pad[2];
};
static_assert(sizeof(InstDev) == 80, "InstDev is 80 bytes");

__device__ inline float3 xform_point(const float4 r0, const float4 r1, const float4 r2, float3 p) {
  return make_float3(((r0.x * p.x + r0.y * p.y) + r0.z * p.z) + r0.w * 1.0f,
                     ((r1.x * p.x + r1.y * p.y) + r1.z * p.z) + r1.w * 1.0f,
                     ((r2.x * p.x + r2.y * p.y) + r2.z * p.z) + r2.w * 1.0f);
}
__device__ inline float3 xform_dir(const float4 r0, const float4 r1, const float4 r2, float3 d) {
  return make_float3(((r0.x * d.x + r0.y * d.y) + r0.z * d.z) + r0.w * 0.0f,
                     ((r1.x * d.x + r1.y * d.y) + r1.z * d.z) + r1.w * 0.0f,
                     ((r2.x * d.x + r2.y * d.y) + r2.z * d.z) + r2.w * 0.0f);
}

// FP16 box coordinate -> FP32 (exact).
__device__ inline float h2f_lo(uint32_t w) {
  return float(__builtin_bit_cast(_Float16, uint16_t(w & 0xFFFFu)));
}
__device__ inline float h2f_hi(uint32_t w) { return float(__builtin_bit_cast(_Float16, uint16_t(w >> 16))); }

// The slab test in the fused form t = fma(b, inv, -(o * inv)) on one FP16
// child record {minx | miny << 16, minz | maxx << 16, maxy | maxz << 16, ref}
// of a BVH4 node.  One rounding per plane instead of two, and each FP16
// bound is a direct operand of v_fma_mix_f32 (widening FP16 to FP32 is
// exact), so no conversion either.
//
// Per axis, one v_perm_b32 puts the bound the ray meets first (the minimum
// where inv >= 0, else the maximum) in the low half and the other in the high
// half, so entry and exit need no min/max: for a box with lo <= hi,
// fma(lo, inv, oi) <= fma(hi, inv, oi) when inv > 0 (rounding is monotonic)
// and the reverse when inv < 0, so the distances equal the min/max form's.
// sel comes from near_far_selectors(inv), once per ray.
//
// Used only by the any-hit traversals of the FP16 nodes, where box tests prune
// and never order the answer: a ray's any-hit answer is whether some triangle
// passes the Moller-Trumbore test, whichever boxes are visited, as long as no
// box holding such a triangle is rejected.  This form and slab round each
// plane distance by about half an ulp of the larger of |b| and |o| (the
// product o * inv, or the difference b - o, is rounded), against boxes
// widened outward by at least one FP16 ulp.  Valid only while fused_ok(oi):
// with a finite -(o * inv) each plane is the exact b * inv - o * inv rounded
// once (no intermediate overflow), so its sign and order are right.
__device__ inline uint3 near_far_selectors(float3 inv) {
  // byte_permute({S0, S1}, sel): bytes 0-3 of S1, then 4-7 of S0
  return make_uint3(inv.x >= 0.0f ? 0x07060100u : 0x01000706u,   // S0 = w.y (maxx hi), S1 = w.x (minx lo)
                    inv.y >= 0.0f ? 0x05040302u : 0x03020504u,   // S0 = w.z (maxy lo), S1 = w.x (miny hi)
                    inv.z >= 0.0f ? 0x07060100u : 0x01000706u);  // S0 = w.z (maxz hi), S1 = w.y (minz lo)
}

__device__ inline float2 slab_fused(float3 oi, float3 inv, uint4 w, uint3 sel) {
  const uint32_t bx = __builtin_amdgcn_perm(w.y, w.x, sel.x);
  const uint32_t by = __builtin_amdgcn_perm(w.z, w.x, sel.y);
  const uint32_t bz = __builtin_amdgcn_perm(w.z, w.y, sel.z);
  const float nx = fmaf(h2f_lo(bx), inv.x, oi.x), fx = fmaf(h2f_hi(bx), inv.x, oi.x);
  const float ny = fmaf(h2f_lo(by), inv.y, oi.y), fy = fmaf(h2f_hi(by), inv.y, oi.y);
  const float nz = fmaf(h2f_lo(bz), inv.z, oi.z), fz = fmaf(h2f_hi(bz), inv.z, oi.z);
  return make_float2(fmaxf(fmaxf(nx, ny), nz), fminf(fminf(fx, fy), fz));
}

// FP32 -> FP16 bits, round to nearest even (v_cvt_f16_f32 in the default mode).
__device__ inline uint32_t f2h_rne(float f) { return __builtin_bit_cast(uint16_t, _Float16(f)); }

// One binary16 step toward -65504 (0xFBFF) and toward +65504 (0x7BFF), the
// two directions cbvh_fat.cpp:447-448 take after rounding: away from the
// target's side a step shrinks the magnitude (the bit pattern minus one),
// toward it the magnitude grows (plus one); either zero steps to the
// smallest subnormal of the target's sign; the target itself stays, and the
// infinity beyond it steps back onto it.  NaN stays.
__device__ inline uint32_t h_step_down(uint32_t h) {
  if ((h & 0x7FFFu) > 0x7C00u || h == 0xFBFFu) return h;
  if (h == 0xFC00u) return 0xFBFFu;
  if ((h & 0x7FFFu) == 0u) return 0x8001u;
  return (h & 0x8000u) ? h + 1u : h - 1u;
}
__device__ inline uint32_t h_step_up(uint32_t h) {
  if ((h & 0x7FFFu) > 0x7C00u || h == 0x7BFFu) return h;
  if (h == 0x7C00u) return 0x7BFFu;
  if ((h & 0x7FFFu) == 0u) return 0x0001u;
  return (h & 0x8000u) ? h - 1u : h + 1u;
}
// cbvh_fat.cpp:447-448: round to nearest, then one step toward -/+ max half.
__device__ inline uint32_t round_to_neg(float v) { return h_step_down(f2h_rne(v)); }
__device__ inline uint32_t round_to_pos(float v) { return h_step_up(f2h_rne(v)); }

// ---------------------------------------------------------------------------
// Builder helpers shared by the single-scene and batched LBVH builders
// ---------------------------------------------------------------------------
constexpr float kLbvhInf = 1e10f;      // lbvh.cpp:53
constexpr float kLbvhEpsBias = 1e-4f;  // lbvh.cpp:54

// 10-bit spread with the magic masks of lbvh.cpp:25-34.
__device__ inline uint32_t spread_bits(int x, int offset) {
  x = (x | (x << 10)) & 0x000F801F;
  x = (x | (x << 4)) & 0x00E181C3;
  x = (x | (x << 2)) & 0x03248649;
  x = (x | (x << 2)) & 0x09249249;
  return uint32_t(x) << offset;
}

// Karras delta: common-prefix length of codes i and j, -1 outside [0, L)
// (lbvh.cpp:386-403).
__device__ inline int kdelta(const uint32_t* __restrict__ c, int L, int i, int j) {
  if (j < 0 || j >= L) return -1;
  const uint32_t x = c[i] ^ c[j];
  return x == 0 ? 32 : int(__builtin_clz(x));
}

// Writes one child box of node pair p (side 0 = left, 1 = right).
__device__ inline void store_box(NodePair* pairs, uint32_t p, uint32_t side, float3 lo, float3 hi) {
  uint32_t* pw = reinterpret_cast<uint32_t*>(pairs + p) + side * 8;
  pw[0] = __float_as_uint(lo.x);
  pw[1] = __float_as_uint(lo.y);
  pw[2] = __float_as_uint(lo.z);
  pw[4] = __float_as_uint(hi.x);
  pw[5] = __float_as_uint(hi.y);
  pw[6] = __float_as_uint(hi.z);
}

}  // namespace hrt
