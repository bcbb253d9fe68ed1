// refit.h -- bottom-up box refit of Karras trees (lbvh.cpp:443-489 computes
// the same boxes with 32 serial sweeps; here each internal node is merged
// once, by the second of its two children to arrive).
//
// Two launches, neither of which needs device-scope ordering inside a kernel:
//
//   k_refit_chunks  one workgroup per chunk of kRefitChunk consecutive leaves.
//                   A node whose leaf range lies inside the chunk only ever
//                   sees arrivals from this workgroup, so its arrival counter
//                   is an LDS word (index = the node's position, which is an
//                   end of its range) and the acquire/release is
//                   workgroup-scoped.  An arrival at a node whose range
//                   crosses the chunk's edge is appended to its tree's list
//                   of top entries instead.
//   k_refit_tops    one workgroup per tree climbs from the top entries; every
//                   node it touches lies above a chunk edge, so all arrivals
//                   at it come from this workgroup too.
//
// On gfx950 a device-scope release writes back the issuing XCD's L2 and the
// matching acquire invalidates it; with thousands of chunk seams those
// flushes, not the arithmetic, would set the refit's time.  The kernel
// boundary between the launches is the only device-wide synchronisation.
//
// The leaf boxes are written into their parents' pairs by the leaf pass
// before these kernels run.  Each thread of k_refit_chunks owns kRefitPer
// consecutive leaves and walks them as a small state machine (arrive, then
// either stop or merge and climb), so lanes stay busy instead of idling
// while a few of them climb.
#pragma once

#include "hrt_device.h"

namespace hrt {

constexpr uint32_t kRefitChunk = 4096;     // leaves per workgroup in k_refit_chunks
constexpr uint32_t kRefitThreads = 256;
constexpr uint32_t kRefitPer = kRefitChunk / kRefitThreads;
constexpr uint32_t kRefitTopsThreads = 512;
constexpr uint32_t kNoNode = 0xFFFFFFFFu;

// Internal nodes on one root-to-leaf path of a Karras tree over unique 30-bit
// codes: each level lengthens the common prefix, so at most 30; a chunk edge
// inside a tree is crossed by at most that many nodes, each of which takes
// at most two entries from below.
constexpr uint32_t kTopsPerEdge = 64;

// Capacity of a tree's top-entry list for a tree of at most n leaves.
__host__ __device__ inline uint32_t refit_tops_capacity(uint64_t n) {
  return uint32_t(kTopsPerEdge * (n / kRefitChunk + 2));
}

// Scene of element i: the last s with off[s] <= i (off holds S + 1 prefix
// sums; empty scenes are skipped because their offset equals the next one).
__device__ inline uint32_t seg_of(const uint32_t* __restrict__ off, uint32_t S, uint32_t i) {
  uint32_t lo = 0, hi = S;  // answer in [lo, hi)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// seg_of for an index that grows with the lane (thread-per-element kernels):
// one search for the wave's first active lane, which the compiler keeps in
// scalar registers and the scalar cache, then a short walk per lane (a
// wave's elements span one or a few segments).
__device__ inline uint32_t seg_of_wave(const uint32_t* __restrict__ off, uint32_t S, uint32_t i) {
  uint32_t s = seg_of(off, S, __builtin_amdgcn_readfirstlane(i));
  while (s + 1 < S && off[s + 1] <= i) s++;
  return s;
}

// A tree's place in the (possibly concatenated) arrays: its index, its first
// and end global leaf, its first pair.  Leaves, parent_leaf and node ranges
// use global leaf numbers; pairs and parent_int are indexed pb + local node.
struct TreeLoc {
  uint32_t s, lb, le, pb;
};

// One tree; the leaf count is on the device (meta[0]).
struct OneTree {
  const uint32_t* meta;
  __device__ uint32_t leaves() const { return meta[0]; }
  __device__ TreeLoc tree(uint32_t) const { return {0u, 0u, meta[0], 0u}; }
  __device__ TreeLoc find(uint32_t) const { return tree(0); }
  __device__ TreeLoc advance(TreeLoc t, uint32_t) const { return t; }
  __device__ uint32_t tops_base(uint32_t) const { return 0u; }
  __device__ uint32_t tops_cap(uint32_t) const { return cap; }
  uint32_t cap;
};

// S trees side by side: leaf_off and tops_off hold S + 1 prefix sums,
// pair_off S + 1 first pairs.
struct ManyTrees {
  const uint32_t* leaf_off;
  const uint32_t* pair_off;
  const uint32_t* tops_off;
  uint32_t S;
  __device__ uint32_t leaves() const { return leaf_off[S]; }
  __device__ TreeLoc tree(uint32_t s) const { return {s, leaf_off[s], leaf_off[s + 1], pair_off[s]}; }
  __device__ TreeLoc find(uint32_t m) const { return tree(seg_of(leaf_off, S, m)); }
  // leaves are walked in increasing order: step to the tree holding m
  __device__ TreeLoc advance(TreeLoc t, uint32_t m) const {
    while (m >= t.le) t = tree(t.s + 1);
    return t;
  }
  __device__ uint32_t tops_base(uint32_t s) const { return tops_off[s]; }
  __device__ uint32_t tops_cap(uint32_t s) const { return tops_off[s + 1] - tops_off[s]; }
};

// Second arrival at pair p: the union of its two child boxes becomes p's box
// in its parent's pair, whose entry (pair << 1 | side) is pp.
__device__ inline void refit_merge_into(NodePair* pairs, uint32_t p, uint32_t pp) {
  const uint32_t* pw = reinterpret_cast<const uint32_t*>(pairs + p);
  const float3 l0 = make_float3(__uint_as_float(pw[0]), __uint_as_float(pw[1]), __uint_as_float(pw[2]));
  const float3 h0 = make_float3(__uint_as_float(pw[4]), __uint_as_float(pw[5]), __uint_as_float(pw[6]));
  const float3 l1 = make_float3(__uint_as_float(pw[8]), __uint_as_float(pw[9]), __uint_as_float(pw[10]));
  const float3 h1 = make_float3(__uint_as_float(pw[12]), __uint_as_float(pw[13]), __uint_as_float(pw[14]));
  const float3 ul = make_float3(fminf(l0.x, l1.x), fminf(l0.y, l1.y), fminf(l0.z, l1.z));
  const float3 uh = make_float3(fmaxf(h0.x, h1.x), fmaxf(h0.y, h1.y), fmaxf(h0.z, h1.z));
  store_box(pairs, pp >> 1, pp & 1u, ul, uh);
}

// Pass 1.  tops: per-tree lists of pair indices (local) to arrive at in
// pass 2; ntops: per-tree counters, zero on entry; err[0] is set when a list
// would overflow (it cannot for trees over unique 30-bit codes).
template <class Trees>
__global__ __launch_bounds__(kRefitThreads) void k_refit_chunks(Trees trees, NodePair* pairs_all,
                                                                const uint32_t* __restrict__ parent_int_all,
                                                                const uint32_t* __restrict__ parent_leaf,
                                                                const uint2* __restrict__ range_all,
                                                                uint32_t* __restrict__ tops,
                                                                uint32_t* __restrict__ ntops,
                                                                uint32_t* __restrict__ err) {
  __shared__ uint32_t arrived[kRefitChunk];
  for (uint32_t k = threadIdx.x; k < kRefitChunk; k += kRefitThreads) arrived[k] = 0u;
  __syncthreads();
  const uint32_t total = trees.leaves();
  const uint32_t chunk_lo = blockIdx.x * kRefitChunk;
  if (chunk_lo >= total) return;
  const uint32_t chunk_hi = min(chunk_lo + kRefitChunk, total) - 1u;
  uint32_t m = chunk_lo + threadIdx.x * kRefitPer;
  const uint32_t m_end = min(m + kRefitPer, chunk_hi + 1u);
  TreeLoc t{};
  if (m < m_end) t = trees.find(m);
  // p: pair (local) of the next arrival; r, pp: its leaf range and its own
  // parent entry, loaded together one step ahead
  uint32_t p = kNoNode, pp = 0u;
  uint2 r = make_uint2(0u, 0u);
  // next leaf of this thread that has a parent (a one-leaf tree has none)
  auto next_leaf = [&]() {
    p = kNoNode;
    for (; m < m_end; m++) {
      t = trees.advance(t, m);
      if (t.le - t.lb > 1u) {
        p = parent_leaf[m] >> 1;
        r = range_all[t.pb + p];
        pp = parent_int_all[t.pb + p];  // unused (and unwritten) for the root
        return;
      }
    }
  };
  next_leaf();
  while (p != kNoNode) {
    bool climb = false;
    if ((r.x < chunk_lo) | (r.y > chunk_hi)) {
      const uint32_t k = __hip_atomic_fetch_add(&ntops[t.s], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (k < trees.tops_cap(t.s)) tops[trees.tops_base(t.s) + k] = p;
      else __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      // releases this thread's box store into pair p, acquires the sibling's
      const uint32_t old =
          __hip_atomic_fetch_add(&arrived[t.lb + p - chunk_lo], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      climb = old != 0u && p != 0u;  // the root's own box is not stored
    }
    if (climb) {
      const uint32_t q = pp >> 1;
      const uint2 rq = range_all[t.pb + q];
      const uint32_t pq = parent_int_all[t.pb + q];
      refit_merge_into(pairs_all + t.pb, p, pp);
      p = q;
      r = rq;
      pp = pq;
    } else {
      m++;
      next_leaf();
    }
  }
}

// Pass 2: one workgroup per tree.  arrive_all: per-pair counters (indexed
// pb + local pair), zero on entry.
template <class Trees>
__global__ __launch_bounds__(kRefitTopsThreads) void k_refit_tops(Trees trees, NodePair* pairs_all,
                                                                  const uint32_t* __restrict__ parent_int_all,
                                                                  const uint32_t* __restrict__ tops,
                                                                  const uint32_t* __restrict__ ntops,
                                                                  uint32_t* arrive_all) {
  const TreeLoc t = trees.tree(blockIdx.x);
  const uint32_t n = min(ntops[t.s], trees.tops_cap(t.s));
  const uint32_t* list = tops + trees.tops_base(t.s);
  NodePair* pairs = pairs_all + t.pb;
  const uint32_t* parent_int = parent_int_all + t.pb;
  uint32_t* arrive = arrive_all + t.pb;
  uint32_t i = threadIdx.x;
  uint32_t p = i < n ? list[i] : kNoNode;
  while (p != kNoNode) {
    const uint32_t old = __hip_atomic_fetch_add(&arrive[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old != 0u && p != 0u) {
      const uint32_t pp = parent_int[p];
      refit_merge_into(pairs, p, pp);
      p = pp >> 1;
    } else {
      i += kRefitTopsThreads;
      p = i < n ? list[i] : kNoNode;
    }
  }
}

}  // namespace hrt
