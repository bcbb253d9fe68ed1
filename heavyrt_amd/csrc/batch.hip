// batch.hip -- many scenes built and traced together in one HIP graph launch.
//
// A batch holds S triangle scenes and S ray sets, concatenated in HBM.  The
// build is the single-scene LBVH pipeline (lbvh.hip; reference
// core/builders/lbvh.cpp:732-827) run as segmented kernels: every stage is
// ONE launch over all scenes, with each thread finding its scene by a binary
// search over the scene offsets, so 256 scenes cost the same number of
// launches as one.  Per scene the results are bit-identical to a solo build:
//   scene box      per-scene reduction over the scene's vertices
//   Morton codes   the scene's box and the N < 256 centroid rule per scene
//   sort           four stable 8-bit passes on the code, then stable passes on
//                  the scene id: the order is (scene, code, triangle id), so
//                  each scene's run of the sorted array is exactly its solo
//                  (code, id) order
//   runs           a run also starts at every scene boundary
//   tree           one bottom-up pass per scene builds the hierarchy and the
//                  boxes together (the same tree as the Karras search); node
//                  pairs, leaf starts and triangle ids are scene-local, as in
//                  a solo build
// The trace is a persistent wave64 nearest-hit kernel over all rays of all
// scenes (ray i finds its scene by binary search); each scene is walked in
// BVH2CommonRT's order (BVH2CommonRT.cpp:49-141), so hits equal solo traces.
//
// The whole pipeline -- fills, build, trace -- is recorded once into a HIP
// graph and replayed with one hipGraphLaunch; nothing in it allocates,
// memsets or synchronises.  The trace kernel's shape (waves per workgroup,
// LDS stack entries per lane, refill threshold, or one ray per lane without
// refill) is chosen among a fixed candidate set by timing each candidate on
// the batch's own rays (hrt_batch_autotune).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "hrt_internal.h"
#include "refit.h"

namespace hrt {
namespace {

constexpr int kB = 256;
constexpr uint32_t kBoxChunk = 4096;   // vertices per box-reduction workgroup

inline uint32_t blocks(uint64_t n) { return uint32_t(std::max<uint64_t>(1, (n + kB - 1) / kB)); }

// Vertex index of corner k of triangle t, clamped into the scene's vertex
// range: the index buffer is filled by the caller on the device, and an
// out-of-range index must not make the build read outside the batch.
__device__ inline uint32_t corner(const uint32_t* __restrict__ idx, uint32_t t, int k, uint32_t nv) {
  return min(idx[3 * t + k], nv - 1);
}


// ---------------------------------------------------------------------------
// build
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kB) void k_b_box_partial(const float4* __restrict__ v, const uint4* __restrict__ chunks,
                                                      float4* __restrict__ part) {
  __shared__ float4 smin[kB];
  __shared__ float4 smax[kB];
  const uint4 c = chunks[blockIdx.x];  // (scene, begin, end, -)
  float4 lo = make_float4(kLbvhInf, kLbvhInf, kLbvhInf, kLbvhInf);
  float4 hi = make_float4(-kLbvhInf, -kLbvhInf, -kLbvhInf, -kLbvhInf);
  for (uint32_t i = c.y + threadIdx.x; i < c.z; i += kB) {
    const float4 p = v[i];
    lo = make_float4(fminf(lo.x, p.x), fminf(lo.y, p.y), fminf(lo.z, p.z), fminf(lo.w, p.w));
    hi = make_float4(fmaxf(hi.x, p.x), fmaxf(hi.y, p.y), fmaxf(hi.z, p.z), fmaxf(hi.w, p.w));
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int s = kB / 2; s > 0; s >>= 1) {
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

// One workgroup per scene: reduce its chunks, derive the Morton scale
// (ZIndexMultCoeff(bboxMin - 1e-4, bboxMax + 1e-4), lbvh.cpp:41-51, 263).
// box12[s] = min xyzw, max xyzw, scale xyzw.
__global__ __launch_bounds__(kB) void k_b_box_final(const float4* __restrict__ part,
                                                    const uint32_t* __restrict__ chunk_off,
                                                    float* __restrict__ box12) {
  __shared__ float4 smin[kB];
  __shared__ float4 smax[kB];
  const uint32_t s = blockIdx.x;
  float4 lo = make_float4(kLbvhInf, kLbvhInf, kLbvhInf, kLbvhInf);
  float4 hi = make_float4(-kLbvhInf, -kLbvhInf, -kLbvhInf, -kLbvhInf);
  for (uint32_t i = chunk_off[s] + threadIdx.x; i < chunk_off[s + 1]; i += kB) {
    const float4 a = part[2 * i], b = part[2 * i + 1];
    lo = make_float4(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z), fminf(lo.w, a.w));
    hi = make_float4(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z), fmaxf(hi.w, b.w));
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int k = kB / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      const float4 a = smin[threadIdx.x + k], b = smax[threadIdx.x + k];
      float4& l = smin[threadIdx.x];
      float4& h = smax[threadIdx.x];
      l = make_float4(fminf(l.x, a.x), fminf(l.y, a.y), fminf(l.z, a.z), fminf(l.w, a.w));
      h = make_float4(fmaxf(h.x, b.x), fmaxf(h.y, b.y), fmaxf(h.z, b.z), fmaxf(h.w, b.w));
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float4 l = smin[0], h = smax[0];
    float* o = box12 + 12 * s;
    o[0] = l.x; o[1] = l.y; o[2] = l.z; o[3] = l.w;
    o[4] = h.x; o[5] = h.y; o[6] = h.z; o[7] = h.w;
    const float ex[4] = {(h.x + kLbvhEpsBias) - (l.x - kLbvhEpsBias), (h.y + kLbvhEpsBias) - (l.y - kLbvhEpsBias),
                         (h.z + kLbvhEpsBias) - (l.z - kLbvhEpsBias), (h.w + kLbvhEpsBias) - (l.w - kLbvhEpsBias)};
    for (int k = 0; k < 4; k++) {
      float inv = 1.0f / ex[k];
      if (k < 3 && fabsf(ex[k]) < 1e-20f) inv = 0.0f;
      o[8 + k] = 1024.0f * inv;
    }
  }
}

struct SceneOffsets {
  const uint32_t* vert;  // S + 1
  const uint32_t* tri;   // S + 1
  const uint32_t* pair;  // S + 1 (max(ntris - 1, 1) pairs reserved per scene)
  const uint32_t* ray;   // S + 1
  const uint32_t* tops;  // S + 1 (each scene's share of the refit's top-entry lists)
  uint32_t S;
};

__global__ __launch_bounds__(kB) void k_b_morton(const float4* __restrict__ v, const uint32_t* __restrict__ idx,
                                                 SceneOffsets off, uint32_t T, const float* __restrict__ box12,
                                                 uint32_t* __restrict__ codes_in, uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * kB + threadIdx.x;
  if (i >= T) return;
  const uint32_t s = seg_of_wave(off.tri, off.S, i);
  const uint32_t n = off.tri[s + 1] - off.tri[s];
  const uint32_t nv = off.vert[s + 1] - off.vert[s];
  const float4* vs = v + off.vert[s];
  const float4 v1 = vs[corner(idx, i, 0, nv)];
  const float4 v2 = vs[corner(idx, i, 1, nv)];
  const float4 v3 = vs[corner(idx, i, 2, nv)];
  float cx, cy, cz;
  if (n < 256) {  // lbvh.cpp:279: centroid for small meshes, box centre otherwise
    cx = 0.33333333f * ((v1.x + v2.x) + v3.x);
    cy = 0.33333333f * ((v1.y + v2.y) + v3.y);
    cz = 0.33333333f * ((v1.z + v2.z) + v3.z);
  } else {
    cx = (fminf(v1.x, fminf(v2.x, v3.x)) + fmaxf(v1.x, fmaxf(v2.x, v3.x))) * 0.5f;
    cy = (fminf(v1.y, fminf(v2.y, v3.y)) + fmaxf(v1.y, fmaxf(v2.y, v3.y))) * 0.5f;
    cz = (fminf(v1.z, fminf(v2.z, v3.z)) + fmaxf(v1.z, fmaxf(v2.z, v3.z))) * 0.5f;
  }
  const float* b = box12 + 12 * s;
  const int ix = int(b[8] * (cx - b[0]));
  const int iy = int(b[9] * (cy - b[1]));
  const int iz = int(b[10] * (cz - b[2]));
  const uint32_t code = spread_bits(ix, 0) | spread_bits(iy, 1) | spread_bits(iz, 2);
  codes_in[i] = code;
  keys[i] = code;
  vals[i] = i;
}

// keys for the scene passes of the sort
__global__ __launch_bounds__(kB) void k_b_scene_keys(const uint32_t* __restrict__ vals, SceneOffsets off, uint32_t T,
                                                     uint32_t* __restrict__ keys) {
  const uint32_t i = blockIdx.x * kB + threadIdx.x;
  if (i >= T) return;
  keys[i] = seg_of(off.tri, off.S, vals[i]);
}

__global__ __launch_bounds__(kB) void k_b_gather_codes(const uint32_t* __restrict__ vals,
                                                       const uint32_t* __restrict__ codes_in, uint32_t T,
                                                       uint32_t* __restrict__ keys) {
  const uint32_t i = blockIdx.x * kB + threadIdx.x;
  if (i >= T) return;
  keys[i] = codes_in[vals[i]];
}

__global__ __launch_bounds__(kB) void k_b_run_flags(const uint32_t* __restrict__ codes, SceneOffsets off,
                                                    uint32_t T, uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * kB + threadIdx.x;
  if (i >= T) return;
  const uint32_t s = seg_of_wave(off.tri, off.S, i);
  flags[i] = (i == off.tri[s] || codes[i - 1] != codes[i]) ? 1u : 0u;
}

// run_start holds scene-local sorted positions; leaf_off[s] = first global leaf
// of scene s, leaf_off[S] = total leaves.
__global__ __launch_bounds__(kB) void k_b_run_compact(const uint32_t* __restrict__ codes,
                                                      const uint32_t* __restrict__ flags,
                                                      const uint32_t* __restrict__ pos, SceneOffsets off, uint32_t T,
                                                      uint32_t* __restrict__ run_start, uint32_t* __restrict__ ucodes,
                                                      uint32_t* __restrict__ leaf_off) {
  const uint32_t i = blockIdx.x * kB + threadIdx.x;
  if (i >= T) return;
  const uint32_t s = seg_of_wave(off.tri, off.S, i);
  if (flags[i]) {
    run_start[pos[i]] = i - off.tri[s];
    ucodes[pos[i]] = codes[i];
  }
  if (i == off.tri[s]) leaf_off[s] = pos[i];
  if (i == T - 1) leaf_off[off.S] = pos[i] + flags[i];
}

// ---------------------------------------------------------------------------
// Hierarchy and boxes in one bottom-up pass per scene: the agglomerative LBVH
// construction of Apetrei (2014, "Fast and Simple Agglomerative LBVH
// Construction").  Over a scene's sorted unique codes c[lb..le), a node
// covering leaves [l, r] has as parent the node that extends it across the
// boundary where the adjacent codes share the longer prefix, i.e. whose xor is
// smaller: across r (the parent splits at r and the node is its left child)
// when l == lb or c[r] ^ c[r+1] < c[l-1] ^ c[l], otherwise across l - 1 (the
// node is the right child).  The two xors never tie for unique sorted codes,
// so this is the binary radix tree that Karras's top-down search builds
// (lbvh.cpp:405-441), and its node numbers follow from the side: a left
// child is node r, a right child node l, the root node 0 (scene-local).
//
// Each split g (between leaves g and g+1) has two slots, one per child side,
// holding half of the parent's pair -- {lo.xyz, child ref, hi.xyz, far end of
// the child's leaf range} -- so the parent's pair is its two slots, the
// range ends zeroed.  A child writes its slot and then arrives at g; the
// second to arrive reads its sibling's slot, writes the parent's pair at the
// parent's own number once it knows the parent's side, and climbs on.
//
// Synchronisation follows refit.h, on three levels, none of which needs
// device scope inside a kernel: a split whose two leaves lie in one
// workgroup's range is arrived at only by that workgroup.
//   k_b_leaves        kB leaves per workgroup: the leaf halves, then every
//                     node inside the block, halves and codes in LDS
//   k_b_climb_chunks  kRefitChunk leaves per workgroup, from the events of
//                     its blocks (arrivals that left a block), LDS exchange
//                     of the child's far end
//   k_b_climb_tops    one workgroup per scene, from the arrivals that left a
//                     chunk, with per-split counters in global memory
// At each level an arrival at a split on the range's edge, and a split left
// with one arrival when the workgroup ends, pass to the next level.
// ---------------------------------------------------------------------------

constexpr uint32_t kClimbDone = 0xFFFFFFFFu;  // exchange word after the second arrival

// Events per leaf-pass block (of kB leaves): arrivals of nodes formed in the
// block at nodes that are not.  Those cross one of the block's two edges; at
// most 30 nodes cross an edge (one per common-prefix length of 30-bit codes),
// each with at most two children from the block: at most 120.
constexpr uint32_t kEventCap = 128;

// True when node [l, r] of a scene over global leaves [lb, le) is its
// parent's left child.  Requires [l, r] to be a proper sub-range.
__device__ inline bool b_is_left(const uint32_t* __restrict__ ucodes, uint32_t lb, uint32_t le, uint32_t l,
                                 uint32_t r) {
  if (l == lb) return true;
  if (r + 1 == le) return false;
  return (ucodes[r] ^ ucodes[r + 1]) < (ucodes[l - 1] ^ ucodes[l]);
}

// Half of a node pair: {lo.xyz, ref}, {hi.xyz, far}.
struct Half {
  uint4 a, b;
};

__device__ inline Half load_half(const uint4* __restrict__ slots, uint32_t k) {
  return Half{slots[2 * size_t(k)], slots[2 * size_t(k) + 1]};
}

__device__ inline void store_half(uint4* __restrict__ slots, uint32_t k, const Half& h) {
  slots[2 * size_t(k)] = h.a;
  slots[2 * size_t(k) + 1] = h.b;
}

// The node's pair from its children's halves (range ends dropped).
__device__ inline void put_pair(NodePair* pairs, uint32_t k, const Half& h0, const Half& h1) {
  uint4* w = reinterpret_cast<uint4*>(pairs + k);
  w[0] = h0.a;
  w[1] = make_uint4(h0.b.x, h0.b.y, h0.b.z, 0u);
  w[2] = h1.a;
  w[3] = make_uint4(h1.b.x, h1.b.y, h1.b.z, 0u);
}

// The node's own half: the union of its children's boxes, ref k, far end f.
__device__ inline Half union_half(const Half& h0, const Half& h1, uint32_t k, uint32_t f) {
  Half u;
  u.a.x = __float_as_uint(fminf(__uint_as_float(h0.a.x), __uint_as_float(h1.a.x)));
  u.a.y = __float_as_uint(fminf(__uint_as_float(h0.a.y), __uint_as_float(h1.a.y)));
  u.a.z = __float_as_uint(fminf(__uint_as_float(h0.a.z), __uint_as_float(h1.a.z)));
  u.a.w = k;
  u.b.x = __float_as_uint(fmaxf(__uint_as_float(h0.b.x), __uint_as_float(h1.b.x)));
  u.b.y = __float_as_uint(fmaxf(__uint_as_float(h0.b.y), __uint_as_float(h1.b.y)));
  u.b.z = __float_as_uint(fmaxf(__uint_as_float(h0.b.z), __uint_as_float(h1.b.z)));
  u.b.w = f;
  return u;
}

// Appends top-list entry e (scene-local split << 1 | side) for scene s.
__device__ inline void push_top(const ManyTrees& trees, uint32_t s, uint32_t e, uint32_t* __restrict__ tops,
                                uint32_t* __restrict__ ntops, uint32_t* __restrict__ err) {
  const uint32_t k = __hip_atomic_fetch_add(&ntops[s], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (k < trees.tops_cap(s)) tops[trees.tops_base(s) + k] = e;
  else __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One thread per global leaf m: its half (box, leaf ref, far end m) starts the
// block's climb, its triangles go into leaf order with scene-local ids, and
// its run length into the scene's longest run.
__global__ __launch_bounds__(kB) void k_b_leaves(const float4* __restrict__ v, const uint32_t* __restrict__ idx,
                                                 const uint32_t* __restrict__ ids_sorted,
                                                 const uint32_t* __restrict__ run_start_all,
                                                 const uint32_t* __restrict__ leaf_off, SceneOffsets off,
                                                 const uint32_t* __restrict__ ucodes,
                                                 NodePair* __restrict__ pairs_all, uint4* __restrict__ slots,
                                                 uint4* __restrict__ ev, uint32_t* __restrict__ evn,
                                                 TriPre* __restrict__ tris, uint32_t* __restrict__ perm,
                                                 uint32_t* __restrict__ max_run, uint32_t* __restrict__ err) {
  __shared__ Half bh[kB][2];         // halves of the block's splits, one per side
  __shared__ uint32_t bother[kB];    // exchange words, as in k_b_climb_chunks
  __shared__ uint32_t bcode[kB + 2];  // codes of leaves blo - 1 .. bhi + 1
  __shared__ uint32_t nev;
  const uint32_t total = leaf_off[off.S];
  const uint32_t blo = blockIdx.x * kB;
  const uint32_t bhi = min(blo + kB, max(total, blo + 1u)) - 1u;  // last leaf of the block (when it has any)
  bother[threadIdx.x] = 0u;
  for (uint32_t k = threadIdx.x; k < kB + 2; k += kB) {
    const uint32_t q = blo + k - 1u;  // wraps for blo = 0, k = 0: not read
    bcode[k] = (k > 0 || blo > 0) && q < total ? ucodes[q] : 0u;
  }
  if (threadIdx.x == 0) nev = 0u;
  __syncthreads();
  const uint32_t m = blockIdx.x * kB + threadIdx.x;
  const bool valid = m < total;
  uint32_t s = 0, run = 0;
  // this leaf's half, and its scene's leaves [lb, lb + L) (L > 1: it climbs)
  Half own{};
  bool two = false;
  uint32_t lb = 0, L = 0;
  if (valid) {
    s = seg_of_wave(leaf_off, off.S, m);
    lb = leaf_off[s];
    L = leaf_off[s + 1] - lb;
    const uint32_t tb = off.tri[s];
    const uint32_t n = off.tri[s + 1] - tb;
    const float4* vs = v + off.vert[s];
    const uint32_t nv = off.vert[s + 1] - off.vert[s];
    const uint32_t st = run_start_all[m];
    const uint32_t e = (m - lb + 1 < L) ? run_start_all[m + 1] : n;
    run = e - st;
    float3 lo = make_float3(kLbvhInf, kLbvhInf, kLbvhInf);
    float3 hi = make_float3(-kLbvhInf, -kLbvhInf, -kLbvhInf);
    for (uint32_t jj = st; jj < e; jj++) {
      const uint32_t t = ids_sorted[tb + jj];
      const float4 a = vs[corner(idx, t, 0, nv)], b = vs[corner(idx, t, 1, nv)], c = vs[corner(idx, t, 2, nv)];
      TriPre r;
      r.a = make_float4(a.x, a.y, a.z, __uint_as_float(t - tb));
      r.e1 = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
      r.e2 = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
      tris[tb + jj] = r;
      perm[tb + jj] = t - tb;
      lo = make_float3(fminf(lo.x, fminf(a.x, fminf(b.x, c.x))), fminf(lo.y, fminf(a.y, fminf(b.y, c.y))),
                       fminf(lo.z, fminf(a.z, fminf(b.z, c.z))));
      hi = make_float3(fmaxf(hi.x, fmaxf(a.x, fmaxf(b.x, c.x))), fmaxf(hi.y, fmaxf(a.y, fmaxf(b.y, c.y))),
                       fmaxf(hi.z, fmaxf(a.z, fmaxf(b.z, c.z))));
    }
    NodePair* pairs = pairs_all + off.pair[s];
    if (L == 1) {
      // root is a leaf: pair 0 = (root leaf, empty dummy), as in lbvh.hip
      store_box(pairs, 0, 0, lo, hi);
      store_box(pairs, 0, 1, make_float3(0.f, 0.f, 0.f), make_float3(0.f, 0.f, 0.f));
      uint32_t* pw = reinterpret_cast<uint32_t*>(pairs);
      pw[3] = pack_leaf(st, e - st);
      pw[7] = 0u;
      pw[11] = 0x80000000u;
      pw[15] = 0u;
    } else {
      two = true;
      own = Half{make_uint4(__float_as_uint(lo.x), __float_as_uint(lo.y), __float_as_uint(lo.z), pack_leaf(st, e - st)),
                 make_uint4(__float_as_uint(hi.x), __float_as_uint(hi.y), __float_as_uint(hi.z), m)};
    }
  }
  // Climb inside the block: nodes whose leaves all lie in [blo, bhi] are
  // formed here, with the halves in LDS.  An arrival at a split on the
  // block's edge, and a split left with one arrival, store the half in its
  // slot and become an event for k_b_climb_chunks.
  if (two) {
    const uint32_t le = lb + L;
    NodePair* pairs = pairs_all + off.pair[s];
    auto side_left = [&](uint32_t l, uint32_t r) {  // b_is_left on the block's codes
      if (l == lb) return true;
      if (r + 1 == le) return false;
      return (bcode[r + 1 - blo] ^ bcode[r + 2 - blo]) < (bcode[l - blo] ^ bcode[l + 1 - blo]);
    };
    uint32_t l = m, r = m;
    bool leaf = true;
    Half h0{}, h1{};
    for (;;) {
      if (!leaf && l == lb && r + 1 == le) {
        put_pair(pairs, 0, h0, h1);
        break;
      }
      const bool left = side_left(l, r);
      const uint32_t g = left ? r : l - 1, side = left ? 0u : 1u, far = left ? l : r;
      if (!leaf) {
        const uint32_t k = (left ? r : l) - lb;
        put_pair(pairs, k, h0, h1);
        own = union_half(h0, h1, k, far);
      }
      if (g < blo || g >= bhi) {
        store_half(slots, 2 * g + side, own);
        const uint32_t i = atomicAdd(&nev, 1u);
        if (i < kEventCap) ev[size_t(blockIdx.x) * kEventCap + i] = make_uint4(g, (far << 1) | side, s, 0u);
        break;
      }
      bh[g - blo][side] = own;
      const uint32_t old = __hip_atomic_exchange(&bother[g - blo], (far + 1u) | (side << 31), __ATOMIC_ACQ_REL,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0u) break;
      bother[g - blo] = kClimbDone;
      const Half sib = bh[g - blo][side ^ 1u];
      if (left) {
        r = (old & 0x7FFFFFFFu) - 1u;
        h0 = own;
        h1 = sib;
      } else {
        l = (old & 0x7FFFFFFFu) - 1u;
        h0 = sib;
        h1 = own;
      }
      leaf = false;
    }
  }
  __syncthreads();
  // splits left with one arrival (the other child crosses the block's edge)
  if (blo < total && threadIdx.x < bhi - blo) {
    const uint32_t v = bother[threadIdx.x];
    if (v != 0u && v != kClimbDone) {
      const uint32_t g = blo + threadIdx.x, side = v >> 31, far = (v & 0x7FFFFFFFu) - 1u;
      store_half(slots, 2 * g + side, bh[threadIdx.x][side]);
      const uint32_t i = atomicAdd(&nev, 1u);
      if (i < kEventCap) ev[size_t(blockIdx.x) * kEventCap + i] = make_uint4(g, (far << 1) | side, s, 0u);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    evn[blockIdx.x] = min(nev, kEventCap);
    if (nev > kEventCap) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // Longest run per scene.  A wave's leaves span one or a few scenes: the
  // lanes of each scene reduce first, then one atomic per scene and wave
  // (a per-leaf atomic would put millions of operations on a few addresses).
  uint64_t todo = __ballot(valid);
  while (todo) {
    const int leader = __builtin_ctzll(todo);
    const uint32_t ls = __shfl(s, leader);
    const bool mine = valid && s == ls;
    uint32_t mv = mine ? run : 0u;
    for (int o = 32; o > 0; o >>= 1) mv = max(mv, __shfl_xor(mv, o));
    if (int(threadIdx.x & 63) == leader) atomicMax(&max_run[ls], mv);
    todo &= ~__ballot(mine);
  }
}

// Chunk pass: one workgroup per kRefitChunk leaves, from the events of the
// chunk's leaf-pass blocks.  other[k]: 0 before any arrival at split
// chunk_lo + k, then the first child's (far end + 1) | side << 31, then
// kClimbDone.
constexpr uint32_t kBlocksPerChunk = kRefitChunk / kB;

__global__ __launch_bounds__(kRefitThreads) void k_b_climb_chunks(ManyTrees trees, const uint32_t* __restrict__ ucodes,
                                                                 uint4* __restrict__ slots, NodePair* pairs_all,
                                                                 const uint4* __restrict__ ev,
                                                                 const uint32_t* __restrict__ evn,
                                                                 uint32_t* __restrict__ tops,
                                                                 uint32_t* __restrict__ ntops,
                                                                 uint32_t* __restrict__ err) {
  __shared__ uint32_t other[kRefitChunk];
  __shared__ uint32_t base[kBlocksPerChunk + 1];
  for (uint32_t k = threadIdx.x; k < kRefitChunk; k += kRefitThreads) other[k] = 0u;
  const uint32_t total = trees.leaves();
  const uint32_t chunk_lo = blockIdx.x * kRefitChunk;
  if (chunk_lo >= total) return;  // the whole workgroup
  const uint32_t chunk_hi = min(chunk_lo + kRefitChunk, total) - 1u;
  const uint32_t b0 = chunk_lo / kB, nb = chunk_hi / kB - b0 + 1u;
  if (threadIdx.x == 0) {
    uint32_t a = 0;
    for (uint32_t k = 0; k < nb; k++) {
      base[k] = a;
      a += evn[b0 + k];
    }
    base[nb] = a;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < base[nb]; i += kRefitThreads) {
    uint32_t k = 0;
    while (base[k + 1] <= i) k++;
    const uint4 e = ev[size_t(b0 + k) * kEventCap + (i - base[k])];
    const TreeLoc t = trees.tree(e.z);
    NodePair* pairs = pairs_all + t.pb;
    // the arriving child: its half is in slot (g, side); [l, r] and the
    // children's halves h0, h1 describe the node formed on a second arrival
    uint32_t g = e.x, side = e.y & 1u, far = e.y >> 1, l = 0, r = 0;
    bool pre = true;
    Half own{}, h0{}, h1{};
    for (;;) {
      if (!pre) {
        if (l == t.lb && r + 1 == t.le) {  // the root
          put_pair(pairs, 0, h0, h1);
          break;
        }
        const bool left = b_is_left(ucodes, t.lb, t.le, l, r);
        g = left ? r : l - 1;
        side = left ? 0u : 1u;
        far = left ? l : r;
        const uint32_t kk = (left ? r : l) - t.lb;
        put_pair(pairs, kk, h0, h1);
        own = union_half(h0, h1, kk, far);
        store_half(slots, 2 * g + side, own);
      }
      if (g < chunk_lo || g >= chunk_hi) {  // the sibling starts outside the chunk
        push_top(trees, t.s, ((g - t.lb) << 1) | side, tops, ntops, err);
        break;
      }
      // releases this thread's slot, acquires the sibling's
      const uint32_t old = __hip_atomic_exchange(&other[g - chunk_lo], (far + 1u) | (side << 31), __ATOMIC_ACQ_REL,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0u) break;
      other[g - chunk_lo] = kClimbDone;
      if (pre) own = load_half(slots, 2 * g + side);
      const Half sib = load_half(slots, 2 * g + (side ^ 1u));
      if (side == 0u) {
        l = far;
        r = (old & 0x7FFFFFFFu) - 1u;
        h0 = own;
        h1 = sib;
      } else {
        l = (old & 0x7FFFFFFFu) - 1u;
        r = far;
        h0 = sib;
        h1 = own;
      }
      pre = false;
    }
  }
  // splits left with one arrival: the other child is finished by the top pass
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < chunk_hi - chunk_lo; k += kRefitThreads) {
    const uint32_t v = other[k];
    if (v != 0u && v != kClimbDone) {
      const uint32_t g = chunk_lo + k;
      const TreeLoc tg = trees.find(g);
      push_top(trees, tg.s, ((g - tg.lb) << 1) | (v >> 31), tops, ntops, err);
    }
  }
}

// Top pass: one workgroup per scene.  arrive_all: per-split counters
// (indexed pb + local split), zero on entry.
__global__ __launch_bounds__(kRefitTopsThreads) void k_b_climb_tops(ManyTrees trees, const uint32_t* __restrict__ ucodes,
                                                                   uint4* slots, NodePair* pairs_all,
                                                                   const uint32_t* __restrict__ tops,
                                                                   const uint32_t* __restrict__ ntops,
                                                                   uint32_t* arrive_all) {
  const TreeLoc t = trees.tree(blockIdx.x);
  const uint32_t n = min(ntops[t.s], trees.tops_cap(t.s));
  const uint32_t* list = tops + trees.tops_base(t.s);
  NodePair* pairs = pairs_all + t.pb;
  uint32_t* arrive = arrive_all + t.pb;
  for (uint32_t i = threadIdx.x; i < n; i += kRefitTopsThreads) {
    uint32_t g = t.lb + (list[i] >> 1), side = list[i] & 1u;
    Half own = load_half(slots, 2 * g + side);
    for (;;) {
      // releases this thread's slot, acquires the sibling's
      const uint32_t old =
          __hip_atomic_fetch_add(&arrive[g - t.lb], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0u) break;
      const Half sib = load_half(slots, 2 * g + (side ^ 1u));
      const Half h0 = side ? sib : own, h1 = side ? own : sib;
      const uint32_t l = h0.b.w, r = h1.b.w;
      if (l == t.lb && r + 1 == t.le) {
        put_pair(pairs, 0, h0, h1);
        break;
      }
      const bool left = b_is_left(ucodes, t.lb, t.le, l, r);
      const uint32_t k = (left ? r : l) - t.lb;
      put_pair(pairs, k, h0, h1);
      g = left ? r : l - 1;
      side = left ? 0u : 1u;
      own = union_half(h0, h1, k, left ? l : r);
      store_half(slots, 2 * g + side, own);
    }
  }
}

// ---------------------------------------------------------------------------
// trace
// ---------------------------------------------------------------------------
// Clears the queue counters of the trace's XcdRayPool (one per XCD).
__global__ void k_b_reset(uint32_t* c) {
  if (threadIdx.x < kXcds) c[threadIdx.x * kXcdQueueStride] = 0u;
}

// Persistent nearest-hit over all scenes' rays.  kWaves waves per workgroup,
// kLds stack entries per lane in LDS (the rest of the 32-entry worst case in
// a per-lane global spill area), refill when kRefill lanes are idle.
template <int kWaves, int kLds, int kRefill, uint32_t kClaim>
__global__ __launch_bounds__(64 * kWaves) void k_b_trace(const NodePair* __restrict__ pairs_all,
                                                         const TriPre* __restrict__ tris_all, SceneOffsets off,
                                                         const Ray* __restrict__ rays, uint32_t R,
                                                         uint2* __restrict__ hits, uint32_t* __restrict__ queue,
                                                         uint32_t* __restrict__ spill) {
  constexpr int kBlk = 64 * kWaves;
  __shared__ uint32_t stack_mem[kLds * kBlk];
  const uint32_t lanes = gridDim.x * kBlk;
  const ShortStack<kLds, kBlk> stk{stack_mem + threadIdx.x, spill + blockIdx.x * kBlk + threadIdx.x, lanes};

  bool active = false;
  uint32_t rid = 0;
  const NodePair* P = pairs_all;
  const TriPre* Tr = tris_all;
  float3 o = make_float3(0.f, 0.f, 0.f), d = o, inv = o;
  float tnear = 0.f;
  bool any = false;
  HitRec h;
  h.t = 0.f;
  h.prim = kInvalid;
  h.u = h.v = 0.f;
  int top = 0;
  uint32_t cur = kInvalid;
  XcdRayPool pool(blockIdx.x);  // ray indices claimed by this wave (its XCD's scenes first)

  while (true) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = __popcll(idle);
    if (!pool.drained() && (nidle >= uint32_t(kRefill) || nidle == 64)) {
      const uint64_t idx = pool.take<kClaim>(nidle, lanes_below(idle), queue, R);
      if (!active) {
        if (idx < R) {
          rid = uint32_t(idx);
          const uint32_t s = seg_of(off.ray, off.S, rid);
          P = pairs_all + off.pair[s];
          Tr = tris_all + off.tri[s];
          const Ray r = rays[rid];
          o = make_float3(r.pos_near.x, r.pos_near.y, r.pos_near.z);
          d = make_float3(r.dir_far.x, r.dir_far.y, r.dir_far.z);
          tnear = r.pos_near.w;
          any = r.dir_far.w <= 0.0f;
          h.t = any ? -r.dir_far.w : r.dir_far.w;
          h.prim = kInvalid;
          h.u = h.v = 0.f;
          inv = safe_inverse(d);
          top = 0;
          cur = 0;
          active = true;
        }
      }
    }
    if (__ballot(active) == 0) break;
    while (true) {
      bool done = false;
      if (active) {
        if ((cur & kLeafBit) == 0) {
          const float4* pp = reinterpret_cast<const float4*>(P + cur);
          const float4 lmin = pp[0], lmax = pp[1], rmin = pp[2], rmax = pp[3];
          const float2 t0 = slab(o, inv, lmin.x, lmin.y, lmin.z, lmax.x, lmax.y, lmax.z);
          const float2 t1 = slab(o, inv, rmin.x, rmin.y, rmin.z, rmax.x, rmax.y, rmax.z);
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
          for (uint32_t k = s; k < s + c; k++) mt_test(o, d, tnear, Tr[k], h);
          cur = (top > 0) ? stk.pop(top) : kInvalid;
        }
        done = (cur == kInvalid) || (any && h.prim != kInvalid);
      }
      if (done) {
        hits[rid] = make_uint2(__float_as_uint(h.t), h.prim);
        active = false;
      }
      const uint32_t nact = __popcll(__ballot(active));
      if (nact == 0 || (!pool.drained() && 64 - nact >= uint32_t(kRefill))) break;
    }
  }
}

// One ray per lane, no refill (kBlk threads per workgroup, full LDS stack).
template <int kBlk>
__global__ __launch_bounds__(kBlk) void k_b_trace_simple(const NodePair* __restrict__ pairs_all,
                                                         const TriPre* __restrict__ tris_all, SceneOffsets off,
                                                         const Ray* __restrict__ rays, uint32_t R,
                                                         uint2* __restrict__ hits) {
  __shared__ uint32_t stack_mem[kStackDepth * kBlk];
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= R) return;
  uint32_t* stack = stack_mem + threadIdx.x;
  const uint32_t sc = seg_of_wave(off.ray, off.S, i);
  const NodePair* P = pairs_all + off.pair[sc];
  const TriPre* Tr = tris_all + off.tri[sc];
  const Ray r = rays[i];
  const float3 o = make_float3(r.pos_near.x, r.pos_near.y, r.pos_near.z);
  const float3 d = make_float3(r.dir_far.x, r.dir_far.y, r.dir_far.z);
  const float tnear = r.pos_near.w;
  const bool any = r.dir_far.w <= 0.0f;
  const float3 inv = safe_inverse(d);
  HitRec h;
  h.t = any ? -r.dir_far.w : r.dir_far.w;
  h.prim = kInvalid;
  h.u = h.v = 0.0f;
  int top = 0;
  uint32_t cur = 0;
  while (true) {
    if ((cur & kLeafBit) == 0) {
      const float4* pp = reinterpret_cast<const float4*>(P + cur);
      const float4 lmin = pp[0], lmax = pp[1], rmin = pp[2], rmax = pp[3];
      const float2 t0 = slab(o, inv, lmin.x, lmin.y, lmin.z, lmax.x, lmax.y, lmax.z);
      const float2 t1 = slab(o, inv, rmin.x, rmin.y, rmin.z, rmax.x, rmax.y, rmax.z);
      const bool h0 = (t0.x <= t0.y) && (t0.y >= tnear) && (t0.x <= h.t);
      const bool h1 = (t1.x <= t1.y) && (t1.y >= tnear) && (t1.x <= h.t);
      const uint32_t c0 = __float_as_uint(lmin.w), c1 = __float_as_uint(rmin.w);
      if (h0 && h1) {
        const bool left_first = t0.x <= t1.x;
        cur = left_first ? c0 : c1;
        stack[top * kBlk] = left_first ? c1 : c0;
        top++;
        continue;
      }
      if (h0 || h1) {
        cur = h0 ? c0 : c1;
        continue;
      }
    } else if (cur != kInvalid) {
      const uint32_t s = leaf_start(cur), c = leaf_count(cur);
      for (uint32_t k = s; k < s + c; k++) mt_test(o, d, tnear, Tr[k], h);
      if (any && h.prim != kInvalid) break;
    }
    if (top == 0) break;
    top--;
    cur = stack[top * kBlk];
  }
  hits[i] = make_uint2(__float_as_uint(h.t), h.prim);
}

// ---------------------------------------------------------------------------
// trace-kernel candidates
// ---------------------------------------------------------------------------
struct Candidate {
  int waves, lds, refill, claim;  // lds == 0: one ray per lane, `waves` * 64 threads
  const void* fn;
};

#define HRT_P(W, S, F, C) {W, S, F, C, reinterpret_cast<const void*>(&k_b_trace<W, S, F, C>)}
#define HRT_S(W) {W, 0, 0, 0, reinterpret_cast<const void*>(&k_b_trace_simple<64 * W>)}
const Candidate kCandidates[] = {
    HRT_P(1, 16, 16, 64),  HRT_P(1, 16, 16, 128), HRT_P(1, 16, 16, 256), HRT_P(1, 16, 32, 64),
    HRT_P(1, 16, 32, 128), HRT_P(1, 16, 8, 128),  HRT_P(1, 8, 16, 128),  HRT_P(1, 24, 16, 128),
    HRT_P(1, 32, 16, 128), HRT_P(2, 16, 16, 128), HRT_P(4, 16, 16, 128), HRT_S(1),
    HRT_S(4),              HRT_P(1, 24, 16, 64),  HRT_P(1, 24, 32, 64),  HRT_P(1, 24, 32, 128),
    HRT_P(1, 24, 32, 32),
};
#undef HRT_P
#undef HRT_S
constexpr int kNumCandidates = int(sizeof(kCandidates) / sizeof(kCandidates[0]));
constexpr int kDefaultCandidate = 1;

int resident_blocks(const void* fn, int threads) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) != hipSuccess || per_cu <= 0) per_cu = 4;
  return cus * per_cu;
}

}  // namespace
}  // namespace hrt

// ===========================================================================
// batch object and C-ABI
// ===========================================================================
struct hrt_batch {
  hrt_ctx* ctx = nullptr;
  uint32_t S = 0;
  uint64_t V = 0, T = 0, R = 0, P = 0;
  std::vector<uint32_t> vert_off, tri_off, ray_off, pair_off, tops_off;
  uint32_t nchunks = 0;
  // device: offsets (5 x (S + 1)), chunk table, chunk offsets
  uint32_t* d_off = nullptr;
  uint4* d_chunks = nullptr;
  uint32_t* d_chunk_off = nullptr;
  // inputs / outputs
  float4* verts = nullptr;
  uint32_t* idx = nullptr;
  hrt::Ray* rays = nullptr;
  uint2* hits = nullptr;
  // trees
  hrt::NodePair* pairs = nullptr;
  hrt::TriPre* tris = nullptr;
  uint32_t* perm = nullptr;
  // scratch: one allocation, carved on the host
  void* scratch = nullptr;
  uint32_t *codes_in, *ka, *va, *kb, *vb, *sort_tmp, *flags, *pos, *scan_tmp, *run_start, *ucodes, *arrive,
      *leaf_off, *max_run, *ntops, *refit_err, *tops, *queue, *spill;
  uint4* slots;  // two halves per split (see k_b_climb_chunks)
  uint4* ev;     // kEventCap events per leaf-pass block, and their counts
  uint32_t* evn;
  float4* part;
  float* box12;
  size_t spill_words = 0;
  // graph
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int candidate = hrt::kDefaultCandidate;
  std::vector<uint32_t> grid;  // per candidate, fixed at creation (no queries during capture)
  std::vector<float> tuning_ms;
  bool built = false;
};

namespace {

using hrt::fail;
using hrt::fail_hip;

hrt::SceneOffsets offsets(const hrt_batch* b) {
  hrt::SceneOffsets o;
  o.vert = b->d_off;
  o.tri = b->d_off + (b->S + 1);
  o.pair = b->d_off + 2 * (b->S + 1);
  o.ray = b->d_off + 3 * (b->S + 1);
  o.tops = b->d_off + 4 * (b->S + 1);
  o.S = b->S;
  return o;
}

void drop_graph(hrt_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Enqueues the build (fills included) on `s`.
hipError_t enqueue_build(hrt_batch* b, hipStream_t s) {
  using namespace hrt;
  const SceneOffsets off = offsets(b);
  const uint32_t T = uint32_t(b->T);
  hipError_t e;
  if ((e = fill_words(b->arrive, b->P, 0u, s)) != hipSuccess) return e;
  // longest runs, top-entry counters and the overflow flag (one region)
  if ((e = fill_words(b->max_run, 2 * uint64_t(b->S) + 1, 0u, s)) != hipSuccess) return e;
  k_b_box_partial<<<b->nchunks, kB, 0, s>>>(b->verts, b->d_chunks, b->part);
  k_b_box_final<<<b->S, kB, 0, s>>>(b->part, b->d_chunk_off, b->box12);
  k_b_morton<<<blocks(T), kB, 0, s>>>(b->verts, b->idx, off, T, b->box12, b->codes_in, b->ka, b->va);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // four passes on the code: the sorted pairs end in (ka, va)
  if ((e = radix_sort(b->ka, b->va, b->kb, b->vb, T, 0, 4, b->sort_tmp, b->refit_err, s)) != hipSuccess) return e;
  uint32_t *kin = b->ka, *vin = b->va;
  // stable passes on the scene id (most significant): the (scene, code, id) order
  if (b->S > 1) {
    int np = 0;
    while ((b->S - 1) >> (8 * np)) np++;
    k_b_scene_keys<<<blocks(T), kB, 0, s>>>(vin, off, T, kin);
    if ((e = radix_sort(b->ka, b->va, b->kb, b->vb, T, 0, np, b->sort_tmp, b->refit_err, s)) != hipSuccess)
      return e;
    if (np & 1) {
      kin = b->kb;
      vin = b->vb;
    }
  }
  // sorted ids in vin; sorted codes gathered into kin
  k_b_gather_codes<<<blocks(T), kB, 0, s>>>(vin, b->codes_in, T, kin);
  k_b_run_flags<<<blocks(T), kB, 0, s>>>(kin, off, T, b->flags);
  if ((e = scan_exclusive(b->flags, b->pos, T, b->scan_tmp, s)) != hipSuccess) return e;
  k_b_run_compact<<<blocks(T), kB, 0, s>>>(kin, b->flags, b->pos, off, T, b->run_start, b->ucodes, b->leaf_off);
  k_b_leaves<<<blocks(T), kB, 0, s>>>(b->verts, b->idx, vin, b->run_start, b->leaf_off, off, b->ucodes, b->pairs,
                                      b->slots, b->ev, b->evn, b->tris, b->perm, b->max_run, b->refit_err);
  const ManyTrees trees{b->leaf_off, off.pair, off.tops, b->S};
  k_b_climb_chunks<<<uint32_t((T + kRefitChunk - 1) / kRefitChunk), kRefitThreads, 0, s>>>(
      trees, b->ucodes, b->slots, b->pairs, b->ev, b->evn, b->tops, b->ntops, b->refit_err);
  k_b_climb_tops<<<b->S, kRefitTopsThreads, 0, s>>>(trees, b->ucodes, b->slots, b->pairs, b->tops, b->ntops,
                                                    b->arrive);
  return hipGetLastError();
}

hipError_t enqueue_trace(hrt_batch* b, int cand, hipStream_t s) {
  using namespace hrt;
  if (b->R == 0) return hipSuccess;
  const SceneOffsets off = offsets(b);
  const Candidate& c = kCandidates[cand];
  const uint32_t R = uint32_t(b->R);
  const int threads = 64 * c.waves;
  if (c.lds == 0) {
    void* args[] = {&b->pairs, &b->tris, const_cast<SceneOffsets*>(&off), &b->rays, const_cast<uint32_t*>(&R),
                    &b->hits};
    return hipLaunchKernel(c.fn, dim3(b->grid[cand]), dim3(threads), args, 0, s);
  }
  k_b_reset<<<1, kXcds, 0, s>>>(b->queue);
  void* args[] = {&b->pairs, &b->tris, const_cast<SceneOffsets*>(&off), &b->rays, const_cast<uint32_t*>(&R),
                  &b->hits, &b->queue, &b->spill};
  return hipLaunchKernel(c.fn, dim3(b->grid[cand]), dim3(threads), args, 0, s);
}

// Workgroups per candidate: one ray per lane for the simple kernels; for the
// persistent ones the resident count, capped so every lane's spill area fits.
void plan_grids(hrt_batch* b) {
  using namespace hrt;
  b->grid.assign(kNumCandidates, 1);
  for (int k = 0; k < kNumCandidates; k++) {
    const Candidate& c = kCandidates[k];
    const uint64_t threads = 64ull * c.waves;
    const uint64_t want = std::max<uint64_t>(1, (b->R + threads - 1) / threads);
    if (c.lds == 0) {
      b->grid[k] = uint32_t(want);
      continue;
    }
    uint64_t g = std::min<uint64_t>(want, uint64_t(resident_blocks(c.fn, int(threads))));
    if (c.lds < kStackDepth) g = std::min<uint64_t>(g, b->spill_words / (threads * uint64_t(kStackDepth - c.lds)));
    b->grid[k] = uint32_t(std::max<uint64_t>(1, g));
  }
}

hrt_status capture_graph(hrt_batch* b) {
  drop_graph(b);
  hipError_t e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return fail_hip(e, "graph capture");
  hipError_t ee = enqueue_build(b, b->cap);
  if (ee == hipSuccess) ee = enqueue_trace(b, b->candidate, b->cap);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(b->cap, &g);
  if (ee != hipSuccess || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail_hip(ee != hipSuccess ? ee : e, "graph capture");
  }
  b->graph = g;
  if ((e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0)) != hipSuccess) {
    drop_graph(b);
    return fail_hip(e, "graph instantiate");
  }
  return HRT_OK;
}

void free_batch(hrt_batch* b) {
  drop_graph(b);
  void* ptrs[] = {b->d_off, b->d_chunks, b->d_chunk_off, b->verts, b->idx, b->rays, b->hits,
                  b->pairs, b->tris,     b->perm,        b->scratch};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (b->cap) (void)hipStreamDestroy(b->cap);
}

}  // namespace

extern "C" {

hrt_status hrt_batch_create(hrt_ctx* ctx, uint32_t nScenes, const uint32_t* nVerts, const uint32_t* nTris,
                            const uint32_t* nRays, hrt_batch** out) {
  using namespace hrt;
  if (!ctx || !out || !nVerts || !nTris || !nRays) return fail(HRT_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (nScenes == 0 || nScenes > 65536) return fail(HRT_ERR_INVALID_ARG, "a batch holds 1 to 65536 scenes");
  hrt_batch* b = new (std::nothrow) hrt_batch();
  if (!b) return fail(HRT_ERR_OUT_OF_MEMORY, "host allocation failed");
  b->ctx = ctx;
  b->S = nScenes;
  const uint32_t S = nScenes;
  b->vert_off.assign(S + 1, 0);
  b->tri_off.assign(S + 1, 0);
  b->ray_off.assign(S + 1, 0);
  b->pair_off.assign(S + 1, 0);
  b->tops_off.assign(S + 1, 0);
  uint64_t V = 0, T = 0, R = 0, P = 0, TOPS = 0;
  std::vector<uint4> chunks;
  std::vector<uint32_t> chunk_off(S + 1, 0);
  for (uint32_t s = 0; s < S; s++) {
    if (nTris[s] == 0 || nVerts[s] == 0) {
      delete b;
      return fail(HRT_ERR_INVALID_ARG, "every scene needs at least one triangle and one vertex");
    }
    if (nTris[s] > 0x00FFFFFFu) {
      delete b;
      return fail(HRT_ERR_INVALID_ARG, "more than 16777215 triangles in one scene (24-bit leaf start)");
    }
    b->vert_off[s] = uint32_t(V);
    b->tri_off[s] = uint32_t(T);
    b->ray_off[s] = uint32_t(R);
    b->pair_off[s] = uint32_t(P);
    b->tops_off[s] = uint32_t(TOPS);
    chunk_off[s] = uint32_t(chunks.size());
    for (uint64_t c = 0; c < nVerts[s]; c += kBoxChunk)
      chunks.push_back(make_uint4(s, uint32_t(V + c), uint32_t(V + std::min<uint64_t>(nVerts[s], c + kBoxChunk)), 0));
    V += nVerts[s];
    T += nTris[s];
    R += nRays[s];
    P += std::max<uint32_t>(nTris[s] > 1 ? nTris[s] - 1 : 1, 1);
    TOPS += refit_tops_capacity(nTris[s]);
  }
  if (V > 0xFFFFFFFFull || T > 0x3FFFFFFFull || R > 0x7FFFFFFFull || P > 0x7FFFFFFFull || TOPS > 0x7FFFFFFFull) {
    delete b;
    return fail(HRT_ERR_INVALID_ARG, "batch too large (at most 2^30-1 triangles, 2^31-1 rays, 2^32-1 vertices)");
  }
  b->vert_off[S] = uint32_t(V);
  b->tri_off[S] = uint32_t(T);
  b->ray_off[S] = uint32_t(R);
  b->pair_off[S] = uint32_t(P);
  b->tops_off[S] = uint32_t(TOPS);
  chunk_off[S] = uint32_t(chunks.size());
  b->V = V;
  b->T = T;
  b->R = R;
  b->P = P;
  b->nchunks = uint32_t(chunks.size());

  hipError_t e = hipSetDevice(ctx->device);
  auto al = [&](auto** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(bytes, 16));
  };
  al(&b->d_off, 4 * 5 * size_t(S + 1));
  al(&b->d_chunks, sizeof(uint4) * chunks.size());
  al(&b->d_chunk_off, 4 * size_t(S + 1));
  al(&b->verts, 16 * V);
  al(&b->idx, 12 * T);
  al(&b->rays, 32 * R);
  al(&b->hits, 8 * R);
  al(&b->pairs, 64 * P);
  al(&b->tris, 48 * T);
  al(&b->perm, 4 * T);
  // scratch words
  const int maxw = 4, minlds = 8;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  b->spill_words = size_t(cus) * 32 * 64 * maxw * (kStackDepth - minlds) / 4;  // enough for every candidate's grid
  size_t words = 0;
  auto take = [&](size_t n) { size_t o = words; words += (n + 63) & ~size_t(63); return o; };
  const size_t o_codes = take(T), o_ka = take(T), o_va = take(T), o_kb = take(T), o_vb = take(T);
  const size_t o_sort = take(radix_sort_scratch_words(T, 4)), o_flags = take(T), o_pos = take(T);
  const size_t o_scan = take(scan_scratch_words(T)), o_rs = take(T), o_uc = take(T);
  const size_t nlb = blocks(T);
  const size_t o_slots = take(16 * T), o_ev = take(4 * kEventCap * nlb), o_evn = take(nlb), o_arr = take(P), o_loff = take(S + 1), o_mrun = take(2 * size_t(S) + 1);
  const size_t o_tops = take(TOPS);
  const size_t o_q = take(64), o_part = take(8 * size_t(b->nchunks)), o_box = take(12 * size_t(S));
  const size_t o_spill = take(b->spill_words);
  al(&b->scratch, 4 * words);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
  if (e != hipSuccess) {
    free_batch(b);
    delete b;
    return fail_hip(e, "batch allocation");
  }
  uint32_t* w = static_cast<uint32_t*>(b->scratch);
  b->codes_in = w + o_codes; b->ka = w + o_ka; b->va = w + o_va; b->kb = w + o_kb; b->vb = w + o_vb;
  b->sort_tmp = w + o_sort; b->flags = w + o_flags; b->pos = w + o_pos; b->scan_tmp = w + o_scan;
  b->run_start = w + o_rs; b->ucodes = w + o_uc; b->slots = reinterpret_cast<uint4*>(w + o_slots);
  b->ev = reinterpret_cast<uint4*>(w + o_ev); b->evn = w + o_evn;
  b->arrive = w + o_arr; b->leaf_off = w + o_loff; b->max_run = w + o_mrun; b->queue = w + o_q;
  b->ntops = b->max_run + S; b->refit_err = b->ntops + S; b->tops = w + o_tops;
  b->part = reinterpret_cast<float4*>(w + o_part);
  b->box12 = reinterpret_cast<float*>(w + o_box);
  b->spill = w + o_spill;
  plan_grids(b);
  std::vector<uint32_t> offs;
  offs.insert(offs.end(), b->vert_off.begin(), b->vert_off.end());
  offs.insert(offs.end(), b->tri_off.begin(), b->tri_off.end());
  offs.insert(offs.end(), b->pair_off.begin(), b->pair_off.end());
  offs.insert(offs.end(), b->ray_off.begin(), b->ray_off.end());
  offs.insert(offs.end(), b->tops_off.begin(), b->tops_off.end());
  e = hipMemcpy(b->d_off, offs.data(), 4 * offs.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_chunks, chunks.data(), sizeof(uint4) * chunks.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_chunk_off, chunk_off.data(), 4 * chunk_off.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    free_batch(b);
    delete b;
    return fail_hip(e, "batch upload");
  }
  *out = b;
  return HRT_OK;
}

void hrt_batch_destroy(hrt_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->cap);
  free_batch(b);
  delete b;
}

hrt_status hrt_batch_buffers(hrt_batch* b, float** verts4, uint32_t** idx, float** rays, uint32_t** hits) {
  if (!b) return fail(HRT_ERR_INVALID_ARG, "batch is null");
  if (verts4) *verts4 = reinterpret_cast<float*>(b->verts);
  if (idx) *idx = b->idx;
  if (rays) *rays = reinterpret_cast<float*>(b->rays);
  if (hits) *hits = reinterpret_cast<uint32_t*>(b->hits);
  return HRT_OK;
}

hrt_status hrt_batch_build(hrt_batch* b, void* stream) {
  if (!b) return fail(HRT_ERR_INVALID_ARG, "batch is null");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : b->ctx->stream;
  hipError_t e = hipSetDevice(b->ctx->device);
  if (e == hipSuccess) e = enqueue_build(b, s);
  if (e != hipSuccess) return fail_hip(e, "batch build");
  b->built = true;
  return HRT_OK;
}

hrt_status hrt_batch_trace(hrt_batch* b, void* stream) {
  if (!b) return fail(HRT_ERR_INVALID_ARG, "batch is null");
  if (!b->built) return fail(HRT_ERR_NOT_COMMITTED, "the batch has not been built (hrt_batch_build or hrt_batch_run)");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : b->ctx->stream;
  hipError_t e = hipSetDevice(b->ctx->device);
  if (e == hipSuccess) e = enqueue_trace(b, b->candidate, s);
  if (e != hipSuccess) return fail_hip(e, "batch trace");
  return HRT_OK;
}

hrt_status hrt_batch_run(hrt_batch* b, void* stream) {
  if (!b) return fail(HRT_ERR_INVALID_ARG, "batch is null");
  hipError_t e = hipSetDevice(b->ctx->device);
  if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
  if (!b->exec) {
    hrt_status st = capture_graph(b);
    if (st != HRT_OK) return st;
  }
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : b->ctx->stream;
  if ((e = hipGraphLaunch(b->exec, s)) != hipSuccess) return fail_hip(e, "graph launch");
  b->built = true;
  return HRT_OK;
}

int hrt_batch_candidate_count(void) { return hrt::kNumCandidates; }

hrt_status hrt_batch_candidate_info(int i, int* wavesPerGroup, int* ldsStackEntries, int* refillAt, int* claim) {
  if (i < 0 || i >= hrt::kNumCandidates) return fail(HRT_ERR_INVALID_ARG, "no such candidate");
  if (wavesPerGroup) *wavesPerGroup = hrt::kCandidates[i].waves;
  if (ldsStackEntries) *ldsStackEntries = hrt::kCandidates[i].lds;
  if (refillAt) *refillAt = hrt::kCandidates[i].refill;
  if (claim) *claim = hrt::kCandidates[i].claim;
  return HRT_OK;
}

hrt_status hrt_batch_set_candidate(hrt_batch* b, int i) {
  if (!b) return fail(HRT_ERR_INVALID_ARG, "batch is null");
  if (i < 0 || i >= hrt::kNumCandidates) return fail(HRT_ERR_INVALID_ARG, "no such candidate");
  if (i != b->candidate) {
    b->candidate = i;
    drop_graph(b);
  }
  return HRT_OK;
}

int hrt_batch_get_candidate(const hrt_batch* b) { return b ? b->candidate : -1; }

hrt_status hrt_batch_autotune(hrt_batch* b, int reps, int* chosen, float* msPerCandidate) {
  if (!b) return fail(HRT_ERR_INVALID_ARG, "batch is null");
  if (reps <= 0) reps = 3;
  hipError_t e = hipSetDevice(b->ctx->device);
  if (e == hipSuccess) e = hipDeviceSynchronize();  // nothing of this batch may still be running
  hipStream_t s = b->cap;
  if (e == hipSuccess && !b->built) {
    e = enqueue_build(b, s);
    b->built = e == hipSuccess;
  }
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&ev0);
  if (e == hipSuccess) e = hipEventCreate(&ev1);
  int best = b->candidate;
  float best_ms = 0.0f;
  b->tuning_ms.assign(hrt::kNumCandidates, 0.0f);
  for (int c = 0; c < hrt::kNumCandidates && e == hipSuccess; c++) {
    e = enqueue_trace(b, c, s);  // warm-up
    float ms_min = 0.0f;
    for (int r = 0; r < reps && e == hipSuccess; r++) {
      e = hipEventRecord(ev0, s);
      if (e == hipSuccess) e = enqueue_trace(b, c, s);
      if (e == hipSuccess) e = hipEventRecord(ev1, s);
      if (e == hipSuccess) e = hipEventSynchronize(ev1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
      if (r == 0 || ms < ms_min) ms_min = ms;
    }
    b->tuning_ms[c] = ms_min;
    if (msPerCandidate) msPerCandidate[c] = ms_min;
    if (c == 0 || ms_min < best_ms) {
      best_ms = ms_min;
      best = c;
    }
  }
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (e != hipSuccess) return fail_hip(e, "autotune");
  hrt_status st = hrt_batch_set_candidate(b, best);
  if (st != HRT_OK) return st;
  if (chosen) *chosen = best;
  return HRT_OK;
}

hrt_status hrt_batch_status(hrt_batch* b, uint32_t* badScene) {
  if (!b) return fail(HRT_ERR_INVALID_ARG, "batch is null");
  if (badScene) *badScene = 0xFFFFFFFFu;
  if (!b->built) return fail(HRT_ERR_NOT_COMMITTED, "the batch has not been built");
  hipError_t e = hipSetDevice(b->ctx->device);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  std::vector<uint32_t> mr(b->S);
  uint32_t rerr = 0;
  if (e == hipSuccess) e = hipMemcpy(mr.data(), b->max_run, 4ull * b->S, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&rerr, b->refit_err, 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail_hip(e, "batch status");
  for (uint32_t s = 0; s < b->S; s++) {
    if (mr[s] > 127) {
      if (badScene) *badScene = s;
      return fail(HRT_ERR_LEAF_OVERFLOW, "scene " + std::to_string(s) + ": " + std::to_string(mr[s]) +
                                             " triangles share one Morton code; the 7-bit leaf count "
                                             "(raytrace_common.h:40) cannot hold them, so its hits are not valid");
    }
  }
  if (rerr) return fail(HRT_ERR_UNSUPPORTED, "a device-side check of the build failed (sort look-back or refit list)");
  return HRT_OK;
}

// Per-scene results after a build: sizes (5 words: triangles, leaves, longest
// run, pairs reserved, first pair), node pairs (16 words each), permutation.
hrt_status hrt_batch_dump(hrt_batch* b, uint32_t scene, int what, void* host_out, size_t bytes) {
  if (!b || !host_out) return fail(HRT_ERR_INVALID_ARG, "batch/out is null");
  if (scene >= b->S) return fail(HRT_ERR_INVALID_ARG, "no such scene");
  if (!b->built) return fail(HRT_ERR_NOT_COMMITTED, "the batch has not been built");
  hipError_t e = hipSetDevice(b->ctx->device);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return fail_hip(e, "dump");
  uint32_t lo[2], mr = 0, rerr = 0;
  if ((e = hipMemcpy(&rerr, b->refit_err, 4, hipMemcpyDeviceToHost)) != hipSuccess) return fail_hip(e, "dump");
  if (rerr) return fail(HRT_ERR_UNSUPPORTED, "a device-side check of the build failed (sort look-back or refit list)");
  if ((e = hipMemcpy(lo, b->leaf_off + scene, 8, hipMemcpyDeviceToHost)) != hipSuccess) return fail_hip(e, "dump");
  if ((e = hipMemcpy(&mr, b->max_run + scene, 4, hipMemcpyDeviceToHost)) != hipSuccess) return fail_hip(e, "dump");
  const uint32_t n = b->tri_off[scene + 1] - b->tri_off[scene];
  const uint32_t L = lo[1] - lo[0];
  const uint32_t np = std::max<uint32_t>(L > 1 ? L - 1 : 1, 1);
  if (what == HRT_DUMP_SIZES) {
    if (bytes < 20) return fail(HRT_ERR_INVALID_ARG, "output buffer too small");
    const uint32_t v[5] = {n, L, mr, np, b->pair_off[scene]};
    std::memcpy(host_out, v, 20);
    return HRT_OK;
  }
  const void* src = nullptr;
  size_t nb = 0;
  if (what == HRT_DUMP_PAIRS) {
    src = b->pairs + b->pair_off[scene];
    nb = 64ull * np;
  } else if (what == HRT_DUMP_PERM) {
    src = b->perm + b->tri_off[scene];
    nb = 4ull * n;
  } else {
    return fail(HRT_ERR_INVALID_ARG, "unknown dump for a batch");
  }
  if (bytes < nb) return fail(HRT_ERR_INVALID_ARG, "output buffer too small");
  e = hipMemcpy(host_out, src, nb, hipMemcpyDeviceToHost);
  return e == hipSuccess ? HRT_OK : fail_hip(e, "dump");
}

}  // extern "C"
