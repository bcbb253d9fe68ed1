// ploc.hip -- the agglomerative hierarchy of the any-hit tree.
//
// An any-hit answer (is there a triangle in (tNear, tFar)?) does not depend on
// the tree, so the tree that any-hit rays walk is this build's own choice
// (no reference counterpart).  Traversal time follows node visits, and a
// Morton split hierarchy is a poor one: this rebuilds the hierarchy above the
// LBVH leaves by parallel locally-ordered clustering (PLOC, Meister & Bittner
// 2018).  The leaves, their order and the leaf-order triangles stay those of
// the LBVH; only the internal nodes change.
//
// Rule (the oracle restates it, and the nodes match it one for one):
//   clusters = the leaves in Morton order;
//   repeat until one cluster is left:
//     each cluster i picks the j != i within [i - R, i + R] minimising the
//     union box's half area (ex*ey + ey*ez + ez*ex, NaN as +inf), ties to
//     the smaller j;
//     every mutual pair i < j becomes internal node (next_id - rank of i
//     among this round's merges), left = i, right = j, at position i; j is
//     dropped; survivors keep their order.
// The pick is a strict order on pairs, so the best pair of a round is always
// mutual and every round merges; ids count down from L - 2, so the root is 0.
//
// One round = nearest neighbours (window in LDS), flags, two scans, merge.
// The host reads two totals per round; the tree is built once per commit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "hrt_internal.h"

namespace hrt {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kPlocRadius = 8;  // window half-width; 4..32 measured, 8 visits fewest on AO rays

// A cluster: box, its reference as a pair child (packed leaf or internal
// node id) in lo.w, its index in the left/right node layout (leaf L-1+m or
// internal id) in hi.w.
struct Cluster {
  float4 lo, hi;
};

__device__ inline float union_half_area(float3 alo, float3 ahi, float3 blo, float3 bhi) {
  const float ex = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
  const float ey = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
  const float ez = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
  const float a = ex * ey + ey * ez + ez * ex;
  return (a == a) ? a : INFINITY;
}

// Leaves (boxes and refs from their parents' pairs) as the first clusters.
__global__ __launch_bounds__(kBlock) void k_ploc_init(const NodePair* __restrict__ pairs,
                                                      const uint32_t* __restrict__ parent_leaf, uint32_t L,
                                                      Cluster* __restrict__ out) {
  const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= L) return;
  const uint32_t pl = parent_leaf[m];
  const uint32_t* pw = reinterpret_cast<const uint32_t*>(pairs + (pl >> 1)) + 8 * (pl & 1u);
  Cluster c;
  c.lo = make_float4(__uint_as_float(pw[0]), __uint_as_float(pw[1]), __uint_as_float(pw[2]), __uint_as_float(pw[3]));
  c.hi = make_float4(__uint_as_float(pw[4]), __uint_as_float(pw[5]), __uint_as_float(pw[6]),
                     __uint_as_float(L - 1 + m));
  out[m] = c;
}

// nn[i] for i < n; the window of the block's clusters is staged in LDS.
__global__ __launch_bounds__(kBlock) void k_ploc_nn(const Cluster* __restrict__ c, uint32_t n,
                                                    uint32_t* __restrict__ nn) {
  constexpr int kR = int(kPlocRadius);
  __shared__ float slo[kBlock + 2 * kR][3];
  __shared__ float shi[kBlock + 2 * kR][3];
  const int64_t first = int64_t(blockIdx.x) * kBlock - kR;
  for (int k = threadIdx.x; k < kBlock + 2 * kR; k += kBlock) {
    const int64_t g = first + k;
    if (g >= 0 && g < int64_t(n)) {
      const Cluster x = c[g];
      slo[k][0] = x.lo.x;
      slo[k][1] = x.lo.y;
      slo[k][2] = x.lo.z;
      shi[k][0] = x.hi.x;
      shi[k][1] = x.hi.y;
      shi[k][2] = x.hi.z;
    }
  }
  __syncthreads();
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int li = threadIdx.x + kR;
  const float3 alo = make_float3(slo[li][0], slo[li][1], slo[li][2]);
  const float3 ahi = make_float3(shi[li][0], shi[li][1], shi[li][2]);
  const uint32_t lo = i > kPlocRadius ? i - kPlocRadius : 0u;
  const uint32_t hi = min(n - 1, i + kPlocRadius);
  float best = 0.0f;
  uint32_t bj = 0xFFFFFFFFu;
  for (uint32_t j = lo; j <= hi; j++) {
    if (j == i) continue;
    const int lj = int(int64_t(j) - first);
    const float a = union_half_area(alo, ahi, make_float3(slo[lj][0], slo[lj][1], slo[lj][2]),
                                    make_float3(shi[lj][0], shi[lj][1], shi[lj][2]));
    if (bj == 0xFFFFFFFFu || a < best) {
      best = a;
      bj = j;
    }
  }
  nn[i] = bj;
}

// merge[i]: i starts a mutual pair; keep[i]: cluster i survives this round
// (as itself or as the new node).
__global__ __launch_bounds__(kBlock) void k_ploc_flags(const uint32_t* __restrict__ nn, uint32_t n,
                                                       uint32_t* __restrict__ merge, uint32_t* __restrict__ keep) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = nn[i];
  const bool mutual = nn[j] == i;
  merge[i] = (mutual && i < j) ? 1u : 0u;
  keep[i] = (mutual && i > j) ? 0u : 1u;
}

__global__ void k_ploc_totals(const uint32_t* __restrict__ merge, const uint32_t* __restrict__ mpos,
                              const uint32_t* __restrict__ keep, const uint32_t* __restrict__ kpos, uint32_t n,
                              uint32_t* __restrict__ out) {
  out[0] = mpos[n - 1] + merge[n - 1];
  out[1] = kpos[n - 1] + keep[n - 1];
}

__global__ __launch_bounds__(kBlock) void k_ploc_merge(const uint32_t* __restrict__ nn,
                                                       const uint32_t* __restrict__ merge,
                                                       const uint32_t* __restrict__ keep,
                                                       const uint32_t* __restrict__ mpos,
                                                       const uint32_t* __restrict__ kpos,
                                                       const Cluster* __restrict__ cin, uint32_t n, uint32_t next_id,
                                                       NodePair* __restrict__ pairs, uint32_t* __restrict__ parent,
                                                       uint32_t* __restrict__ lr, Cluster* __restrict__ cout) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (merge[i]) {
    const Cluster a = cin[i], b = cin[nn[i]];
    const uint32_t id = next_id - mpos[i];
    store_box(pairs, id, 0, make_float3(a.lo.x, a.lo.y, a.lo.z), make_float3(a.hi.x, a.hi.y, a.hi.z));
    store_box(pairs, id, 1, make_float3(b.lo.x, b.lo.y, b.lo.z), make_float3(b.hi.x, b.hi.y, b.hi.z));
    uint32_t* pw = reinterpret_cast<uint32_t*>(pairs + id);
    const uint32_t ra = __float_as_uint(a.lo.w), rb = __float_as_uint(b.lo.w);
    pw[3] = ra;
    pw[7] = 0u;
    pw[11] = rb;
    pw[15] = 0u;
    if ((ra & kLeafBit) == 0u) parent[ra] = id << 1;
    if ((rb & kLeafBit) == 0u) parent[rb] = (id << 1) | 1u;
    if (lr) {
      lr[2 * id] = __float_as_uint(a.hi.w);
      lr[2 * id + 1] = __float_as_uint(b.hi.w);
    }
    Cluster u;
    u.lo = make_float4(fminf(a.lo.x, b.lo.x), fminf(a.lo.y, b.lo.y), fminf(a.lo.z, b.lo.z), __uint_as_float(id));
    u.hi = make_float4(fmaxf(a.hi.x, b.hi.x), fmaxf(a.hi.y, b.hi.y), fmaxf(a.hi.z, b.hi.z), __uint_as_float(id));
    cout[kpos[i]] = u;
  } else if (keep[i]) {
    cout[kpos[i]] = cin[i];
  }
}

// Depth of every internal node (root 0 has none above it); max into out[0].
// A chain longer than 64 steps (impossible for a tree) reports 64.
__global__ __launch_bounds__(kBlock) void k_ploc_depth(const uint32_t* __restrict__ parent, uint32_t ninternal,
                                                       uint32_t* __restrict__ out) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= ninternal) return;
  uint32_t d = 0, p = k;
  while (p != 0u && d < 64u) {
    p = parent[p] >> 1;
    d++;
  }
  atomicMax(out, d);
}

inline uint32_t blocks_for(uint64_t n) { return uint32_t(std::max<uint64_t>(1, (n + kBlock - 1) / kBlock)); }

}  // namespace

hipError_t blas_ploc(Blas& b, hipStream_t s, bool* replaced) {
  *replaced = false;
  const uint32_t L = b.leaves;
  if (L < 3) return hipSuccess;  // one or two leaves: nothing to cluster
  const uint32_t ni = L - 1;
  // device scratch: new pairs and parents, two cluster arrays, nn, flags,
  // positions, scan temp, totals
  const size_t w_scan = scan_scratch_words(L);
  size_t words = 0;
  auto take = [&](size_t n) { size_t o = words; words += (n + 63) & ~size_t(63); return o; };
  const size_t o_ca = take(8ull * L), o_cb = take(8ull * L), o_nn = take(L), o_m = take(L), o_k = take(L);
  const size_t o_mp = take(L), o_kp = take(L), o_scan = take(w_scan), o_tot = take(4);
  NodePair* npairs = nullptr;
  uint32_t *nparent = nullptr, *nlr = nullptr;
  uint32_t* base = nullptr;
  // the same capacities as blas_alloc's (pairs for every triangle but one),
  // so the swapped arrays can stand in for the originals everywhere
  const size_t cap = std::max<size_t>(b.ntris > 1 ? b.ntris - 1 : 1, 1);
  hipError_t e = hipMalloc(&npairs, sizeof(NodePair) * cap);
  if (e == hipSuccess) e = hipMalloc(&nparent, 4ull * cap);
  if (e == hipSuccess && b.karras_lr) e = hipMalloc(&nlr, 8ull * cap);
  if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void**>(&base), 4 * words, s);
  bool ok = false;
  if (e == hipSuccess) {
    Cluster* ca = reinterpret_cast<Cluster*>(base + o_ca);
    Cluster* cb = reinterpret_cast<Cluster*>(base + o_cb);
    uint32_t *nn = base + o_nn, *mf = base + o_m, *kf = base + o_k, *mp = base + o_mp, *kp = base + o_kp;
    uint32_t *stmp = base + o_scan, *tot = base + o_tot;
    k_ploc_init<<<blocks_for(L), kBlock, 0, s>>>(b.pairs, b.parent_leaf, L, ca);
    uint32_t n = L, next_id = L - 2;
    do {
      k_ploc_nn<<<blocks_for(n), kBlock, 0, s>>>(ca, n, nn);
      k_ploc_flags<<<blocks_for(n), kBlock, 0, s>>>(nn, n, mf, kf);
      if ((e = scan_exclusive(mf, mp, n, stmp, s)) != hipSuccess) break;
      if ((e = scan_exclusive(kf, kp, n, stmp, s)) != hipSuccess) break;
      k_ploc_totals<<<1, 1, 0, s>>>(mf, mp, kf, kp, n, tot);
      k_ploc_merge<<<blocks_for(n), kBlock, 0, s>>>(nn, mf, kf, mp, kp, ca, n, next_id, npairs, nparent, nlr,
                                                    cb);
      uint32_t ht[2];
      if ((e = hipMemcpyAsync(ht, tot, 8, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
      if ((e = hipStreamSynchronize(s)) != hipSuccess) break;
      // every round merges at least one pair (see the rule above)
      if (ht[0] == 0 || ht[0] > next_id + 1 || ht[1] != n - ht[0]) {
        e = hipErrorIllegalState;
        break;
      }
      next_id -= ht[0];
      n = ht[1];
      std::swap(ca, cb);
    } while (n > 1);
    if (e == hipSuccess) {
      // the BVH4 traversal stack holds kStackDepth4 entries: at most
      // kMaxLevels4 internal levels
      const uint32_t zero = 0u;
      uint32_t depth = 0;
      e = hipMemcpyAsync(tot, &zero, 4, hipMemcpyHostToDevice, s);
      if (e == hipSuccess) {
        k_ploc_depth<<<blocks_for(ni), kBlock, 0, s>>>(nparent, ni, tot);
        e = hipMemcpyAsync(&depth, tot, 4, hipMemcpyDeviceToHost, s);
      }
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      ok = e == hipSuccess && depth + 1 <= uint32_t(kMaxLevels4);
    }
  }
  if (base) {
    const hipError_t fe = hipFreeAsync(base, s);
    if (e == hipSuccess) e = fe;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess && ok) {
    std::swap(b.pairs, npairs);
    std::swap(b.parent, nparent);
    if (nlr) std::swap(b.karras_lr, nlr);
    b.clustered = true;
    *replaced = true;
  }
  if (npairs) (void)hipFree(npairs);
  if (nparent) (void)hipFree(nparent);
  if (nlr) (void)hipFree(nlr);
  return e;
}

}  // namespace hrt
