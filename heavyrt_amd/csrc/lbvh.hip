// lbvh.hip -- LBVH builder for gfx950.
//
// Stages (reference: core/builders/lbvh.cpp:732-827, BuildFromTriangles):
//   1. scene box over all vertices            lbvh.cpp:228-239, 746
//   2. 30-bit Morton codes per triangle       lbvh.cpp:25-51, 259-287
//   3. stable LSD radix sort of (code, id)    lbvh.cpp:755 (order fixed to (code, id)); sort.hip
//   4. unique-code runs (flag, scan, compact) lbvh.cpp:289-319, 762-767
//   5. Karras 2012 hierarchy                  lbvh.cpp:386-441
//   6. leaves, with the triangles re-laid out lbvh.cpp:350-384,
//      in leaf order                          BVH2CommonRT_host.cpp:45-64
//   7. atomic bottom-up refit (refit.h)       lbvh.cpp:443-489
//
// Every integer result (codes, order, runs, children) is bit-identical to the
// reference's; boxes are min/max unions, exact in any order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

#include "hrt_internal.h"
#include "refit.h"

namespace hrt {
namespace {

constexpr int kBlock = 256;

// Scan tile: 256 threads x 8 items.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

#define HRT_CHECK(x)                         \
  do {                                       \
    hipError_t e_ = (x);                     \
    if (e_ != hipSuccess) return e_;         \
  } while (0)

// ---------------------------------------------------------------------------
// Primitive sources.  The same pipeline builds a BLAS over triangles
// (BuildFromTriangles, lbvh.cpp:732-827) and a TLAS over instance boxes
// (BuildFromBoxes, lbvh.cpp:829-917); a source says what the scene box is
// reduced over, where a primitive's Morton point lies and what its box is.
// ---------------------------------------------------------------------------
struct TriSource {
  static constexpr bool kTris = true;  // the leaf pass also writes leaf-order triangles
  const float4* v;
  const uint32_t* idx;
  uint32_t nverts;
  // scene box over every vertex (lbvh.cpp:228-239)
  __device__ uint32_t box_items() const { return nverts; }
  __device__ void box_item(uint32_t i, float4& lo, float4& hi) const { lo = hi = v[i]; }
  // centroid for small meshes, box centre otherwise (lbvh.cpp:266-283)
  __device__ float3 centre(uint32_t i, uint32_t n) const {
    const float4 v1 = v[idx[3 * i + 0]];
    const float4 v2 = v[idx[3 * i + 1]];
    const float4 v3 = v[idx[3 * i + 2]];
    if (n < 256)
      return make_float3(0.33333333f * ((v1.x + v2.x) + v3.x), 0.33333333f * ((v1.y + v2.y) + v3.y),
                         0.33333333f * ((v1.z + v2.z) + v3.z));
    return make_float3((fminf(v1.x, fminf(v2.x, v3.x)) + fmaxf(v1.x, fmaxf(v2.x, v3.x))) * 0.5f,
                       (fminf(v1.y, fminf(v2.y, v3.y)) + fmaxf(v1.y, fmaxf(v2.y, v3.y))) * 0.5f,
                       (fminf(v1.z, fminf(v2.z, v3.z)) + fmaxf(v1.z, fmaxf(v2.z, v3.z))) * 0.5f);
  }
  __device__ void corners(uint32_t t, float4& a, float4& b, float4& c) const {
    a = v[idx[3 * t + 0]];
    b = v[idx[3 * t + 1]];
    c = v[idx[3 * t + 2]];
  }
  __device__ void prim_box(uint32_t t, float3& lo, float3& hi) const {
    const float4 a = v[idx[3 * t + 0]], b = v[idx[3 * t + 1]], c = v[idx[3 * t + 2]];
    lo = make_float3(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z)));
    hi = make_float3(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z)));
  }
};

struct BoxSource {
  static constexpr bool kTris = false;
  const float4* boxes;  // 2 per primitive: min, max
  uint32_t n;
  // scene box over the primitive boxes (lbvh.cpp:190-203)
  __device__ uint32_t box_items() const { return n; }
  __device__ void box_item(uint32_t i, float4& lo, float4& hi) const {
    lo = boxes[2 * i];
    hi = boxes[2 * i + 1];
  }
  // box centre (lbvh.cpp:251)
  __device__ float3 centre(uint32_t i, uint32_t) const {
    const float4 lo = boxes[2 * i], hi = boxes[2 * i + 1];
    return make_float3((lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f);
  }
  __device__ void prim_box(uint32_t t, float3& lo, float3& hi) const {
    const float4 a = boxes[2 * t], b = boxes[2 * t + 1];
    lo = make_float3(a.x, a.y, a.z);
    hi = make_float3(b.x, b.y, b.z);
  }
};

// ---------------------------------------------------------------------------
// 1. scene box
// ---------------------------------------------------------------------------
template <class Src>
__global__ __launch_bounds__(kBlock) void k_box_partial(Src src, float4* __restrict__ part) {
  __shared__ float4 smin[kBlock];
  __shared__ float4 smax[kBlock];
  float4 lo = make_float4(kLbvhInf, kLbvhInf, kLbvhInf, kLbvhInf);
  float4 hi = make_float4(-kLbvhInf, -kLbvhInf, -kLbvhInf, -kLbvhInf);
  const uint32_t n = src.box_items();
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    float4 a, b;
    src.box_item(i, a, b);
    lo = make_float4(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z), fminf(lo.w, a.w));
    hi = make_float4(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z), fmaxf(hi.w, b.w));
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const float4 a = smin[threadIdx.x + s], b = smax[threadIdx.x + s];
      float4& l = smin[threadIdx.x];
      float4& h = smax[threadIdx.x];
      l = make_float4(fminf(l.x, a.x), fminf(l.y, a.y), fminf(l.z, a.z), fminf(l.w, a.w));
      h = make_float4(fmaxf(h.x, b.x), fmaxf(h.y, b.y), fmaxf(h.z, b.z), fmaxf(h.w, b.w));
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x + 0] = smin[0];
    part[2 * blockIdx.x + 1] = smax[0];
  }
}

// Reduces the partial boxes and derives the Morton scale
// (ZIndexMultCoeff(bboxMin - 1e-4, bboxMax + 1e-4), lbvh.cpp:41-51, 263).
// cubic = 1: the same 1024-cell scale on every axis, set by the largest
// extent (cells are cubes).  The reference scales each axis to its own
// extent; on flat scenes that spends as many code bits on the thin axis as
// on the wide ones.  Used only for the any-hit tree, whose answers do not
// depend on the tree.
__global__ __launch_bounds__(kBlock) void k_box_final(const float4* __restrict__ part, int nparts,
                                                      float* __restrict__ box8, float* __restrict__ zscale4,
                                                      int cubic) {
  __shared__ float4 smin[kBlock];
  __shared__ float4 smax[kBlock];
  float4 lo = make_float4(kLbvhInf, kLbvhInf, kLbvhInf, kLbvhInf);
  float4 hi = make_float4(-kLbvhInf, -kLbvhInf, -kLbvhInf, -kLbvhInf);
  for (int i = threadIdx.x; i < nparts; i += kBlock) {
    const float4 a = part[2 * i], b = part[2 * i + 1];
    lo = make_float4(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z), fminf(lo.w, a.w));
    hi = make_float4(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z), fmaxf(hi.w, b.w));
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const float4 a = smin[threadIdx.x + s], b = smax[threadIdx.x + s];
      float4& l = smin[threadIdx.x];
      float4& h = smax[threadIdx.x];
      l = make_float4(fminf(l.x, a.x), fminf(l.y, a.y), fminf(l.z, a.z), fminf(l.w, a.w));
      h = make_float4(fmaxf(h.x, b.x), fmaxf(h.y, b.y), fmaxf(h.z, b.z), fmaxf(h.w, b.w));
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float4 l = smin[0], h = smax[0];
    box8[0] = l.x; box8[1] = l.y; box8[2] = l.z; box8[3] = l.w;
    box8[4] = h.x; box8[5] = h.y; box8[6] = h.z; box8[7] = h.w;
    const float ex[4] = {(h.x + kLbvhEpsBias) - (l.x - kLbvhEpsBias), (h.y + kLbvhEpsBias) - (l.y - kLbvhEpsBias),
                         (h.z + kLbvhEpsBias) - (l.z - kLbvhEpsBias), (h.w + kLbvhEpsBias) - (l.w - kLbvhEpsBias)};
    if (cubic) {
      const float m = fmaxf(ex[0], fmaxf(ex[1], ex[2]));
      const float inv = (fabsf(m) < 1e-20f) ? 0.0f : 1.0f / m;
      zscale4[0] = zscale4[1] = zscale4[2] = 1024.0f * inv;
      zscale4[3] = 1024.0f * (1.0f / ex[3]);
      return;
    }
    for (int k = 0; k < 4; k++) {
      float inv = 1.0f / ex[k];
      if (k < 3 && fabsf(ex[k]) < 1e-20f) inv = 0.0f;
      zscale4[k] = 1024.0f * inv;
    }
  }
}

// ---------------------------------------------------------------------------
// 2. Morton codes
// ---------------------------------------------------------------------------

template <class Src>
__global__ __launch_bounds__(kBlock) void k_morton(Src src, uint32_t ntris, const float* __restrict__ box8,
                                                   const float* __restrict__ zscale4, uint32_t* __restrict__ codes,
                                                   uint32_t* __restrict__ ids) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= ntris) return;
  const float3 c = src.centre(i, ntris);
  const float sx = zscale4[0] * (c.x - box8[0]);
  const float sy = zscale4[1] * (c.y - box8[1]);
  const float sz = zscale4[2] * (c.z - box8[2]);
  // float -> int truncates toward zero (v_cvt_i32_f32 saturates, NaN -> 0)
  const int ix = int(sx);
  const int iy = int(sy);
  const int iz = int(sz);
  codes[i] = spread_bits(ix, 0) | spread_bits(iy, 1) | spread_bits(iz, 2);
  ids[i] = i;
}

// ---------------------------------------------------------------------------
// 3. stable LSD radix sort: sort.hip
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// exclusive scan of u32 (three phases, recursive over block sums)
// ---------------------------------------------------------------------------
__device__ inline uint32_t block_exclusive_scan(uint32_t x, uint32_t* total) {
  __shared__ uint32_t s[kBlock];
  s[threadIdx.x] = x;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    const uint32_t y = (threadIdx.x >= off) ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += y;
    __syncthreads();
  }
  const uint32_t incl = s[threadIdx.x];
  *total = s[kBlock - 1];
  __syncthreads();
  return incl - x;
}

__global__ __launch_bounds__(kBlock) void k_scan_tile(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                      uint64_t n, uint32_t* __restrict__ sums) {
  const uint64_t base = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * kScanItems;
  uint32_t v[kScanItems];
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    acc += v[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(acc, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0 && sums) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* __restrict__ out, uint64_t n,
                                                     const uint32_t* __restrict__ sums) {
  const uint64_t base = uint64_t(blockIdx.x) * kScanTile;
  const uint32_t add = sums[blockIdx.x];
  for (int k = threadIdx.x; k < kScanTile; k += kBlock)
    if (base + k < n) out[base + k] += add;
}

size_t scan_temp_words(uint64_t n) {
  size_t words = 0;
  while (n > uint64_t(kScanTile)) {
    n = (n + kScanTile - 1) / kScanTile;
    words += n;
  }
  return words + 1;
}

hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* temp, hipStream_t s) {
  const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles <= 1) {
    k_scan_tile<<<1, kBlock, 0, s>>>(in, out, n, nullptr);
    return hipGetLastError();
  }
  uint32_t* sums = temp;
  k_scan_tile<<<dim3(uint32_t(tiles)), kBlock, 0, s>>>(in, out, n, sums);
  HRT_CHECK(hipGetLastError());
  HRT_CHECK(exclusive_scan(sums, sums, tiles, temp + tiles, s));
  k_scan_add<<<dim3(uint32_t(tiles)), kBlock, 0, s>>>(out, n, sums);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// 4. unique-code runs
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_run_flags(const uint32_t* __restrict__ codes, uint32_t n,
                                                      uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  flags[i] = (i == 0 || codes[i - 1] != codes[i]) ? 1u : 0u;
}

// meta[0] = leaves, meta[1] = max run length
__global__ __launch_bounds__(kBlock) void k_run_compact(const uint32_t* __restrict__ codes,
                                                        const uint32_t* __restrict__ flags,
                                                        const uint32_t* __restrict__ pos, uint32_t n,
                                                        uint32_t* __restrict__ run_start,
                                                        uint32_t* __restrict__ ucodes, uint32_t* __restrict__ meta) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (flags[i]) {
    run_start[pos[i]] = i;
    ucodes[pos[i]] = codes[i];
  }
  if (i == n - 1) meta[0] = pos[i] + flags[i];
}

// ---------------------------------------------------------------------------
// 5. Karras 2012
// ---------------------------------------------------------------------------

__device__ inline uint32_t leaf_ref(const uint32_t* __restrict__ run_start, uint32_t L, uint32_t ntris, uint32_t m) {
  const uint32_t s = run_start[m];
  const uint32_t e = (m + 1 < L) ? run_start[m + 1] : ntris;
  return pack_leaf(s, e - s);
}

__global__ __launch_bounds__(kBlock) void k_karras(const uint32_t* __restrict__ ucodes,
                                                   const uint32_t* __restrict__ run_start,
                                                   const uint32_t* __restrict__ meta, uint32_t ntris,
                                                   NodePair* __restrict__ pairs, uint32_t* __restrict__ parent_int,
                                                   uint32_t* __restrict__ parent_leaf, uint2* __restrict__ range,
                                                   uint32_t* __restrict__ karras_lr) {
  const int L = int(meta[0]);
  const int i = int(blockIdx.x * kBlock + threadIdx.x);
  if (i >= L - 1) return;
  const int dr = kdelta(ucodes, L, i, i + 1) - kdelta(ucodes, L, i, i - 1);
  const int d = (dr > 0) - (dr < 0);
  const int dmin = kdelta(ucodes, L, i, i - d);
  int lmax = 2;
  while (kdelta(ucodes, L, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t > 0; t >>= 1)
    if (kdelta(ucodes, L, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = kdelta(ucodes, L, i, j);
  int s = 0;
  for (int t = (l + 1) >> 1; t > 0; t = (t == 1 ? 0 : (t + 1) >> 1))
    if (kdelta(ucodes, L, i, i + (s + t) * d) > dnode) s += t;
  const int gamma = i + s * d + min(d, 0);
  const bool left_leaf = (min(i, j) == gamma);
  const bool right_leaf = (max(i, j) == gamma + 1);
  uint32_t* pw = reinterpret_cast<uint32_t*>(pairs + i);
  uint32_t lref, rref;
  if (left_leaf) {
    lref = leaf_ref(run_start, uint32_t(L), ntris, uint32_t(gamma));
    parent_leaf[gamma] = uint32_t(i) << 1;
  } else {
    lref = uint32_t(gamma);
    parent_int[gamma] = uint32_t(i) << 1;
  }
  if (right_leaf) {
    rref = leaf_ref(run_start, uint32_t(L), ntris, uint32_t(gamma + 1));
    parent_leaf[gamma + 1] = (uint32_t(i) << 1) | 1u;
  } else {
    rref = uint32_t(gamma + 1);
    parent_int[gamma + 1] = (uint32_t(i) << 1) | 1u;
  }
  pw[3] = lref;
  pw[7] = 0u;
  pw[11] = rref;
  pw[15] = 0u;
  range[i] = make_uint2(uint32_t(min(i, j)), uint32_t(max(i, j)));
  if (karras_lr) {
    karras_lr[2 * i + 0] = left_leaf ? uint32_t(L - 1 + gamma) : uint32_t(gamma);
    karras_lr[2 * i + 1] = right_leaf ? uint32_t(L - 1 + gamma + 1) : uint32_t(gamma + 1);
  }
}

// ---------------------------------------------------------------------------
// 6. leaves (and leaf-order triangles)
// ---------------------------------------------------------------------------

// One thread per leaf m: its box, the union of its run's primitive boxes
// (lbvh.cpp:350-384), goes into its parent's pair, where the refit starts.
// For triangles the same gathered vertices also give the leaf-order records
// {A, origId}, {B - A}, {C - A} (BVH2CommonRT_host.cpp:45-64,
// BVH2CommonRT.cpp:26-27) and the permutation (perm may be null: a refit
// keeps it).  meta[1] gets the longest run.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_leaves(Src src, const uint32_t* __restrict__ ids_sorted,
                                                   const uint32_t* __restrict__ run_start,
                                                   uint32_t* __restrict__ meta, uint32_t ntris,
                                                   NodePair* __restrict__ pairs,
                                                   const uint32_t* __restrict__ parent_leaf,
                                                   TriPre* __restrict__ tris, uint32_t* __restrict__ perm) {
  const uint32_t L = meta[0];
  const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= L) return;
  const uint32_t s = run_start[m];
  const uint32_t e = (m + 1 < L) ? run_start[m + 1] : ntris;
  float3 lo = make_float3(kLbvhInf, kLbvhInf, kLbvhInf);
  float3 hi = make_float3(-kLbvhInf, -kLbvhInf, -kLbvhInf);
  for (uint32_t j = s; j < e; j++) {
    const uint32_t t = ids_sorted[j];
    float3 a, b;
    if constexpr (Src::kTris) {
      float4 va, vb, vc;
      src.corners(t, va, vb, vc);
      TriPre r;
      r.a = make_float4(va.x, va.y, va.z, __uint_as_float(t));
      r.e1 = make_float4(vb.x - va.x, vb.y - va.y, vb.z - va.z, 0.0f);
      r.e2 = make_float4(vc.x - va.x, vc.y - va.y, vc.z - va.z, 0.0f);
      tris[j] = r;
      if (perm) perm[j] = t;
      a = make_float3(fminf(va.x, fminf(vb.x, vc.x)), fminf(va.y, fminf(vb.y, vc.y)), fminf(va.z, fminf(vb.z, vc.z)));
      b = make_float3(fmaxf(va.x, fmaxf(vb.x, vc.x)), fmaxf(va.y, fmaxf(vb.y, vc.y)), fmaxf(va.z, fmaxf(vb.z, vc.z)));
    } else {
      src.prim_box(t, a, b);
    }
    lo = make_float3(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z));
    hi = make_float3(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z));
  }
  atomicMax(&meta[1], e - s);  // one address: combined across the wave
  if (L == 1) {
    // Root is a leaf: pair 0 = (root leaf, empty dummy), as the reference's
    // peach layout stores it (lbvh.cpp:497-504, cbvh.h:28-36).
    store_box(pairs, 0, 0, lo, hi);
    store_box(pairs, 0, 1, make_float3(0.f, 0.f, 0.f), make_float3(0.f, 0.f, 0.f));
    uint32_t* pw = reinterpret_cast<uint32_t*>(pairs);
    pw[3] = pack_leaf(s, e - s);
    pw[7] = 0u;
    pw[11] = 0x80000000u;
    pw[15] = 0u;
    return;
  }
  const uint32_t pl = parent_leaf[m];
  store_box(pairs, pl >> 1, pl & 1u, lo, hi);
}

// ---------------------------------------------------------------------------
// BVH4 collapse
// ---------------------------------------------------------------------------
// flag[k] = 1 when internal node k has even depth (it becomes a BVH4 node)
__global__ __launch_bounds__(kBlock) void k_even_depth(const uint32_t* __restrict__ parent_int, uint32_t ninternal,
                                                       uint32_t* __restrict__ flag) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= ninternal) return;
  uint32_t d = 0, p = k;
  while (p != 0) {  // at most 30 steps: the tree has at most 30 internal levels
    p = parent_int[p] >> 1;
    d++;
  }
  flag[k] = (d & 1u) ? 0u : 1u;
}

struct Child {
  float3 lo, hi;
  uint32_t ref;
};

__device__ inline Child pair_child(const NodePair* pairs, uint32_t k, uint32_t side) {
  const uint32_t* pw = reinterpret_cast<const uint32_t*>(pairs + k) + 8 * side;
  Child c;
  c.lo = make_float3(__uint_as_float(pw[0]), __uint_as_float(pw[1]), __uint_as_float(pw[2]));
  c.hi = make_float3(__uint_as_float(pw[4]), __uint_as_float(pw[5]), __uint_as_float(pw[6]));
  c.ref = pw[3];
  return c;
}

// One BVH4 node per even-depth internal node k: k's children, each internal
// child replaced by its two children, left to right
// (cbvh_core.cpp:263-352); internal references become quad indices.
__global__ __launch_bounds__(kBlock) void k_collapse4(const NodePair* __restrict__ pairs,
                                                      const uint32_t* __restrict__ flag,
                                                      const uint32_t* __restrict__ qidx, uint32_t ninternal,
                                                      Quad32* __restrict__ q32, Quad16* __restrict__ q16) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  const bool single = (ninternal == 0);  // the root is a leaf: one node from pair 0
  if (single ? k != 0 : (k >= ninternal || !flag[k])) return;
  Child c[4];
  int n = 0;
  for (uint32_t side = 0; side < 2; side++) {
    const Child ch = pair_child(pairs, k, side);
    if (!single && (ch.ref & kLeafBit) == 0) {
      c[n++] = pair_child(pairs, ch.ref, 0);
      c[n++] = pair_child(pairs, ch.ref, 1);
    } else {
      c[n++] = ch;
    }
  }
  for (; n < 4; n++) {
    c[n].lo = make_float3(1e9f, 1e9f, 1e9f);
    c[n].hi = make_float3(-1e9f, -1e9f, -1e9f);
    c[n].ref = pack_leaf(0, 0);
  }
  uint32_t ref[4];
  for (int i = 0; i < 4; i++) ref[i] = (c[i].ref & kLeafBit) ? c[i].ref : qidx[c[i].ref];
  const uint32_t q = single ? 0u : qidx[k];
  Quad32 o;
  o.lox = make_float4(c[0].lo.x, c[1].lo.x, c[2].lo.x, c[3].lo.x);
  o.hix = make_float4(c[0].hi.x, c[1].hi.x, c[2].hi.x, c[3].hi.x);
  o.loy = make_float4(c[0].lo.y, c[1].lo.y, c[2].lo.y, c[3].lo.y);
  o.hiy = make_float4(c[0].hi.y, c[1].hi.y, c[2].hi.y, c[3].hi.y);
  o.loz = make_float4(c[0].lo.z, c[1].lo.z, c[2].lo.z, c[3].lo.z);
  o.hiz = make_float4(c[0].hi.z, c[1].hi.z, c[2].hi.z, c[3].hi.z);
  o.ref = make_uint4(ref[0], ref[1], ref[2], ref[3]);
  o.pad = make_uint4(0u, 0u, 0u, 0u);
  q32[q] = o;
  Quad16 h;
  for (int i = 0; i < 4; i++) {  // PackNode, cbvh_fat.cpp:477-497
    h.c[i] = make_uint4(round_to_neg(c[i].lo.x) | (round_to_neg(c[i].lo.y) << 16),
                        round_to_neg(c[i].lo.z) | (round_to_pos(c[i].hi.x) << 16),
                        round_to_pos(c[i].hi.y) | (round_to_pos(c[i].hi.z) << 16), ref[i]);
  }
  q16[q] = h;
}

// Leaf-order triangle j with its absolute vertices for the watertight test:
// {A, original id}, {B, 0}, {C, 0} (the vertices exactly as stored, so
// triangles that share a vertex see the same values).
__global__ __launch_bounds__(kBlock) void k_wt_tris(const float4* __restrict__ verts, const uint32_t* __restrict__ idx,
                                                    const uint32_t* __restrict__ perm, uint32_t ntris,
                                                    TriPre* __restrict__ out) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= ntris) return;
  const uint32_t t = perm[j];
  const float4 A = verts[idx[3 * size_t(t) + 0]];
  const float4 B = verts[idx[3 * size_t(t) + 1]];
  const float4 C = verts[idx[3 * size_t(t) + 2]];
  TriPre r;
  r.a = make_float4(A.x, A.y, A.z, __uint_as_float(t));
  r.e1 = make_float4(B.x, B.y, B.z, 0.0f);
  r.e2 = make_float4(C.x, C.y, C.z, 0.0f);
  out[j] = r;
}

inline uint32_t blocks_for(uint64_t n) { return uint32_t((n + kBlock - 1) / kBlock); }

}  // namespace

hipError_t blas_alloc(Blas& b, uint32_t nverts, uint32_t ntris, bool keep_debug) {
  b.nverts = nverts;
  b.ntris = ntris;
  b.keep_debug = keep_debug;
  const size_t npairs = std::max<size_t>(ntris > 1 ? ntris - 1 : 1, 1);
  HRT_CHECK(hipMalloc(&b.verts, sizeof(float4) * std::max<size_t>(nverts, 1)));
  HRT_CHECK(hipMalloc(&b.idx, sizeof(uint32_t) * 3 * size_t(ntris)));
  HRT_CHECK(hipMalloc(&b.pairs, sizeof(NodePair) * npairs));
  HRT_CHECK(hipMalloc(&b.tris, sizeof(TriPre) * size_t(ntris)));
  HRT_CHECK(hipMalloc(&b.perm, sizeof(uint32_t) * size_t(ntris)));
  HRT_CHECK(hipMalloc(&b.parent, sizeof(uint32_t) * npairs));
  HRT_CHECK(hipMalloc(&b.leaf_run, sizeof(uint32_t) * size_t(ntris)));
  HRT_CHECK(hipMalloc(&b.parent_leaf, sizeof(uint32_t) * size_t(ntris)));
  HRT_CHECK(hipMalloc(&b.node_range, sizeof(uint2) * npairs));
  HRT_CHECK(hipMalloc(&b.meta, sizeof(uint32_t) * 64));
  if (keep_debug) {
    HRT_CHECK(hipMalloc(&b.codes_in, sizeof(uint32_t) * size_t(ntris)));
    HRT_CHECK(hipMalloc(&b.codes_sorted, sizeof(uint32_t) * size_t(ntris)));
    HRT_CHECK(hipMalloc(&b.ids_sorted, sizeof(uint32_t) * size_t(ntris)));
    HRT_CHECK(hipMalloc(&b.run_start, sizeof(uint32_t) * size_t(ntris)));
    HRT_CHECK(hipMalloc(&b.ucodes, sizeof(uint32_t) * size_t(ntris)));
    HRT_CHECK(hipMalloc(&b.karras_lr, sizeof(uint32_t) * 2 * npairs));
  }
  return hipSuccess;
}

void blas_free(Blas& b) {
  if (b.any_tree) {
    blas_free(*b.any_tree);
    delete b.any_tree;
    b.any_tree = nullptr;
  }
  void* ptrs[] = {b.leaf_run,   b.parent_leaf, b.node_range, b.meta,       b.tris_wt,
                  b.verts,      b.idx,        b.pairs,      b.tris,      b.perm,   b.parent,
                  b.quads32,    b.quads16,    b.codes_in,   b.codes_sorted, b.ids_sorted,
                  b.run_start,  b.ucodes,     b.karras_lr};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  b = Blas();
}

namespace {

// Fills n words with v (used instead of hipMemsetAsync so the build can be
// recorded into a HIP graph).
__global__ __launch_bounds__(kBlock) void k_fill(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) p[i] = v;
}

inline uint32_t fill_blocks(uint64_t n) { return uint32_t(std::min<uint64_t>(1024, std::max<uint64_t>(1, (n + kBlock - 1) / kBlock))); }

// The LBVH pipeline over any primitive source.  `finish(ids_sorted)` enqueues
// the leaf-order pass; the stream is synchronised before returning and the
// leaf count, longest run and scene box are host values afterwards.
template <class Src, class Finish>
hipError_t lbvh_build(Blas& b, const Src& src, uint32_t box_items, hipStream_t s, Finish&& finish, int cubic = 0) {
  const uint64_t N = b.ntris;
  if (N == 0) return hipErrorInvalidValue;
  const int box_parts = int(std::min<uint64_t>(1024, std::max<uint64_t>(1, (box_items + kBlock - 1) / kBlock)));
  // scratch layout (u32 words)
  const size_t w_codes = N, w_sort = radix_sort_scratch_words(N, 4), w_scan = scan_temp_words(N);
  size_t words = 0;
  auto take = [&](size_t n) { size_t o = words; words += (n + 63) & ~size_t(63); return o; };
  const size_t o_ka = take(w_codes), o_va = take(w_codes), o_kb = take(w_codes), o_vb = take(w_codes);
  const size_t o_sort = take(w_sort), o_scan = take(w_scan);
  const size_t o_flags = take(N), o_pos = take(N), o_uc = take(N), o_arr = take(N);
  const uint32_t tops_cap = refit_tops_capacity(N);
  const size_t o_tops = take(tops_cap);
  const size_t o_part = take(size_t(box_parts) * 8), o_box = take(16);
  uint32_t* base = nullptr;
  HRT_CHECK(hipMallocAsync(reinterpret_cast<void**>(&base), words * 4, s));
  uint32_t *ka = base + o_ka, *va = base + o_va, *kb = base + o_kb, *vb = base + o_vb;
  uint32_t *sort_tmp = base + o_sort, *stmp = base + o_scan;
  uint32_t *flags = base + o_flags, *pos = base + o_pos, *uc = base + o_uc, *arrive = base + o_arr;
  uint32_t* tops = base + o_tops;
  // runs, leaf parents, node ranges and the leaf count stay with the tree (refit)
  uint32_t *rs = b.leaf_run, *pleaf = b.parent_leaf, *meta = b.meta;
  uint2* range = b.node_range;
  float4* part = reinterpret_cast<float4*>(base + o_part);
  float* box = reinterpret_cast<float*>(base + o_box);

  hipError_t err = hipSuccess;
  do {
    // meta: [0] leaves, [1] longest run, [2] refit top entries, [3] refit
    // overflow, [4] sort look-back timeout; flags stay zero unless a check fails
    k_fill<<<1, kBlock, 0, s>>>(meta, 64, 0u);
    k_box_partial<Src><<<box_parts, kBlock, 0, s>>>(src, part);
    k_box_final<<<1, kBlock, 0, s>>>(part, box_parts, box, box + 8, cubic);
    k_morton<Src><<<blocks_for(N), kBlock, 0, s>>>(src, uint32_t(N), box, box + 8, ka, va);
    if ((err = hipGetLastError()) != hipSuccess) break;
    if (b.keep_debug)
      if ((err = hipMemcpyAsync(b.codes_in, ka, 4 * N, hipMemcpyDeviceToDevice, s)) != hipSuccess) break;
    // four 8-bit passes: the sorted pairs end in (ka, va)
    if ((err = radix_sort(ka, va, kb, vb, N, 0, 4, sort_tmp, meta + 4, s)) != hipSuccess) break;
    uint32_t *kin = ka, *vin = va;
    k_run_flags<<<blocks_for(N), kBlock, 0, s>>>(kin, uint32_t(N), flags);
    if ((err = exclusive_scan(flags, pos, N, stmp, s)) != hipSuccess) break;
    k_run_compact<<<blocks_for(N), kBlock, 0, s>>>(kin, flags, pos, uint32_t(N), rs, uc, meta);
    k_fill<<<fill_blocks(N), kBlock, 0, s>>>(arrive, N, 0u);
    k_karras<<<blocks_for(N), kBlock, 0, s>>>(uc, rs, meta, uint32_t(N), b.pairs, b.parent, pleaf, range,
                                              b.karras_lr);
    k_leaves<Src><<<blocks_for(N), kBlock, 0, s>>>(src, vin, rs, meta, uint32_t(N), b.pairs, pleaf, b.tris,
                                                   b.perm);
    // meta[2] counts the refit's top entries, meta[3] flags an overflow (both zeroed above)
    const OneTree tree{meta, tops_cap};
    k_refit_chunks<OneTree><<<uint32_t((N + kRefitChunk - 1) / kRefitChunk), kRefitThreads, 0, s>>>(
        tree, b.pairs, b.parent, pleaf, range, tops, meta + 2, meta + 3);
    k_refit_tops<OneTree><<<1, kRefitTopsThreads, 0, s>>>(tree, b.pairs, b.parent, tops, meta + 2, arrive);
    finish(vin);
    if ((err = hipGetLastError()) != hipSuccess) break;
    if (b.keep_debug) {
      if ((err = hipMemcpyAsync(b.codes_sorted, kin, 4 * N, hipMemcpyDeviceToDevice, s)) != hipSuccess) break;
      if ((err = hipMemcpyAsync(b.ids_sorted, vin, 4 * N, hipMemcpyDeviceToDevice, s)) != hipSuccess) break;
      if ((err = hipMemcpyAsync(b.run_start, rs, 4 * N, hipMemcpyDeviceToDevice, s)) != hipSuccess) break;
      if ((err = hipMemcpyAsync(b.ucodes, uc, 4 * N, hipMemcpyDeviceToDevice, s)) != hipSuccess) break;
    }
    uint32_t hmeta[5];
    float hbox[12];
    if ((err = hipMemcpyAsync(hmeta, meta, sizeof(hmeta), hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    if ((err = hipMemcpyAsync(hbox, box, sizeof(hbox), hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    if ((err = hipStreamSynchronize(s)) != hipSuccess) break;
    if (hmeta[3] != 0u || hmeta[4] != 0u) {  // a device-side check failed: the tree is incomplete
      err = hipErrorIllegalState;
      break;
    }
    b.leaves = hmeta[0];
    b.max_run = hmeta[1];
    std::memcpy(b.box, hbox, sizeof(b.box));
    std::memcpy(b.zscale, hbox + 8, sizeof(b.zscale));
  } while (false);
  hipError_t ferr = hipFreeAsync(base, s);
  if (err == hipSuccess) err = ferr;
  if (err == hipSuccess) err = hipStreamSynchronize(s);
  return err;
}

// Instance records in leaf order.
__global__ __launch_bounds__(kBlock) void k_inst_order(const InstDev* __restrict__ by_id,
                                                       const uint32_t* __restrict__ ids_sorted, uint32_t n,
                                                       InstDev* __restrict__ leaf_order, uint32_t* __restrict__ perm) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const uint32_t t = ids_sorted[j];
  leaf_order[j] = by_id[t];
  perm[j] = t;
}

}  // namespace

hipError_t blas_build(Blas& b, hipStream_t s, bool cubic) {
  const TriSource src{b.verts, b.idx, b.nverts};
  // the leaf pass already wrote the leaf-order triangles
  return lbvh_build(b, src, b.nverts, s, [](const uint32_t*) {}, cubic ? 1 : 0);
}

hipError_t blas_refit(Blas& b, hipStream_t s) {
  const uint64_t N = b.ntris;
  if (N == 0 || !b.leaf_run) return hipErrorInvalidValue;
  const TriSource src{b.verts, b.idx, b.nverts};
  const int box_parts = int(std::min<uint64_t>(1024, std::max<uint64_t>(1, (b.nverts + kBlock - 1) / kBlock)));
  uint32_t* base = nullptr;
  const uint32_t tops_cap = refit_tops_capacity(N);
  const size_t o_tops = (N + 63) & ~size_t(63);
  const size_t o_part = o_tops + ((tops_cap + 63) & ~size_t(63));
  HRT_CHECK(hipMallocAsync(reinterpret_cast<void**>(&base), 4 * (o_part + 8 * size_t(box_parts) + 16), s));
  uint32_t* arrive = base;
  uint32_t* tops = base + o_tops;
  float4* part = reinterpret_cast<float4*>(base + o_part);
  float* box = reinterpret_cast<float*>(base + o_part + 8 * size_t(box_parts));
  k_box_partial<TriSource><<<box_parts, kBlock, 0, s>>>(src, part);
  k_box_final<<<1, kBlock, 0, s>>>(part, box_parts, box, box + 8, 0);
  k_fill<<<fill_blocks(N), kBlock, 0, s>>>(arrive, N, 0u);
  k_fill<<<1, kBlock, 0, s>>>(b.meta + 2, 2, 0u);  // top-entry count, overflow flag
  // the kept permutation gives the leaf order; it is not rewritten
  k_leaves<TriSource><<<blocks_for(N), kBlock, 0, s>>>(src, b.perm, b.leaf_run, b.meta, uint32_t(N), b.pairs,
                                                       b.parent_leaf, b.tris, nullptr);
  const OneTree tree{b.meta, tops_cap};
  k_refit_chunks<OneTree><<<uint32_t((N + kRefitChunk - 1) / kRefitChunk), kRefitThreads, 0, s>>>(
      tree, b.pairs, b.parent, b.parent_leaf, b.node_range, tops, b.meta + 2, b.meta + 3);
  k_refit_tops<OneTree><<<1, kRefitTopsThreads, 0, s>>>(tree, b.pairs, b.parent, tops, b.meta + 2, arrive);
  hipError_t err = hipGetLastError();
  float hbox[8];
  uint32_t hmeta[4];
  if (err == hipSuccess) err = hipMemcpyAsync(hbox, box, sizeof(hbox), hipMemcpyDeviceToHost, s);
  if (err == hipSuccess) err = hipMemcpyAsync(hmeta, b.meta, sizeof(hmeta), hipMemcpyDeviceToHost, s);
  hipError_t ferr = hipFreeAsync(base, s);
  if (err == hipSuccess) err = ferr;
  if (err == hipSuccess) err = hipStreamSynchronize(s);
  if (err == hipSuccess && hmeta[3] != 0u) err = hipErrorIllegalState;  // a refit list overflowed
  if (err != hipSuccess) return err;
  std::memcpy(b.box, hbox, sizeof(b.box));
  // boxes changed: the BVH4 nodes, the watertight records and the any-hit
  // tree are made again on the next commit
  if (b.tris_wt) (void)hipFree(b.tris_wt);
  b.tris_wt = nullptr;
  if (b.quads32) (void)hipFree(b.quads32);
  if (b.quads16) (void)hipFree(b.quads16);
  b.quads32 = nullptr;
  b.quads16 = nullptr;
  b.nquads = 0;
  if (b.any_tree) {
    blas_free(*b.any_tree);
    delete b.any_tree;
    b.any_tree = nullptr;
  }
  return hipSuccess;
}

hipError_t blas_build_any_tree(Blas& b, hipStream_t s) {
  if (b.any_tree) return hipSuccess;
  Blas* a = new (std::nothrow) Blas();
  if (!a) return hipErrorOutOfMemory;
  hipError_t e = blas_alloc(*a, b.nverts, b.ntris, b.keep_debug);
  if (e == hipSuccess) e = hipMemcpyAsync(a->verts, b.verts, 16ull * b.nverts, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(a->idx, b.idx, 12ull * b.ntris, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = blas_build(*a, s, true);
  // a run longer than the 7-bit leaf count cannot be stored: keep the
  // reference tree for any-hit queries then
  bool replaced = false;
  if (e == hipSuccess && a->max_run <= 127) e = blas_ploc(*a, s, &replaced);
  if (e == hipSuccess && a->max_run <= 127) e = blas_build_bvh4(*a, s);
  if (e != hipSuccess || a->max_run > 127) {
    blas_free(*a);
    delete a;
    return e;
  }
  b.any_tree = a;
  return hipSuccess;
}

hipError_t tlas_build(Blas& t, const InstDev* by_id, InstDev* leaf_order, hipStream_t s) {
  const BoxSource src{t.verts, t.ntris};
  const uint32_t N = t.ntris;
  return lbvh_build(t, src, N, s, [&](const uint32_t* ids_sorted) {
    k_inst_order<<<blocks_for(N), kBlock, 0, s>>>(by_id, ids_sorted, N, leaf_order, t.perm);
  });
}

hipError_t blas_build_wt_tris(Blas& b, hipStream_t s) {
  if (b.tris_wt) return hipSuccess;
  HRT_CHECK(hipMalloc(&b.tris_wt, sizeof(TriPre) * size_t(b.ntris)));
  k_wt_tris<<<blocks_for(b.ntris), kBlock, 0, s>>>(b.verts, b.idx, b.perm, b.ntris, b.tris_wt);
  return hipGetLastError();
}

hipError_t blas_build_bvh4(Blas& b, hipStream_t s) {
  if (b.quads32) return hipSuccess;
  const uint32_t ninternal = b.leaves > 1 ? b.leaves - 1 : 0;
  const size_t nflag = std::max<uint32_t>(ninternal, 1);
  const size_t w_scan = scan_temp_words(nflag);
  uint32_t* base = nullptr;
  HRT_CHECK(hipMallocAsync(reinterpret_cast<void**>(&base), 4 * (2 * nflag + w_scan + 64), s));
  uint32_t *flag = base, *qidx = base + nflag, *stmp = base + 2 * nflag;
  hipError_t err = hipSuccess;
  do {
    uint32_t nq = 1;
    if (ninternal > 0) {
      k_even_depth<<<blocks_for(ninternal), kBlock, 0, s>>>(b.parent, ninternal, flag);
      if ((err = exclusive_scan(flag, qidx, ninternal, stmp, s)) != hipSuccess) break;
      uint32_t last[2];
      if ((err = hipMemcpyAsync(&last[0], qidx + ninternal - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
      if ((err = hipMemcpyAsync(&last[1], flag + ninternal - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
      if ((err = hipStreamSynchronize(s)) != hipSuccess) break;
      nq = last[0] + last[1];
    }
    if ((err = hipMalloc(&b.quads32, sizeof(Quad32) * nq)) != hipSuccess) break;
    if ((err = hipMalloc(&b.quads16, sizeof(Quad16) * nq)) != hipSuccess) break;
    k_collapse4<<<blocks_for(std::max<uint32_t>(ninternal, 1)), kBlock, 0, s>>>(b.pairs, flag, qidx, ninternal,
                                                                               b.quads32, b.quads16);
    if ((err = hipGetLastError()) != hipSuccess) break;
    b.nquads = nq;
  } while (false);
  hipError_t ferr = hipFreeAsync(base, s);
  if (err == hipSuccess) err = ferr;
  if (err == hipSuccess) err = hipStreamSynchronize(s);
  if (err != hipSuccess) {
    if (b.quads32) (void)hipFree(b.quads32);
    if (b.quads16) (void)hipFree(b.quads16);
    b.quads32 = nullptr;
    b.quads16 = nullptr;
    b.nquads = 0;
  }
  return err;
}

}  // namespace hrt

// ---------------------------------------------------------------------------
// Building blocks shared with the batched builder (all graph-capturable: no
// allocation, no memset, no synchronisation).
// ---------------------------------------------------------------------------
namespace hrt {

size_t scan_scratch_words(uint64_t n) { return scan_temp_words(n); }

hipError_t scan_exclusive(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* temp, hipStream_t s) {
  return exclusive_scan(in, out, n, temp, s);
}

hipError_t fill_words(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_fill<<<fill_blocks(n), kBlock, 0, s>>>(p, n, v);
  return hipGetLastError();
}

}  // namespace hrt
