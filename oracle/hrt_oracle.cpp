// hrt_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's ray-query path (HeavyRT), used as the
// parity checker for the HIP engine and as the CPU baseline in bench.py.
// Nothing in the shipped library links, loads or calls this file.
//
// What is restated, and where it comes from (reference paths are relative to
// the HeavyRT source tree):
//   Morton codes          core/builders/lbvh.cpp:25-51, 259-287
//   scene box             core/builders/lbvh.cpp:228-239 (all vertices, +-1e10 init)
//   sort                  core/builders/lbvh.cpp:755 -- the reference uses an
//                         unstable std::sort on the code; this restatement
//                         fixes the order to (code, index), i.e. a stable sort
//   unique-code runs      core/builders/lbvh.cpp:289-319, 765
//   leaves                core/builders/lbvh.cpp:350-384
//   Karras 2012           core/builders/lbvh.cpp:386-441
//   refit                 core/builders/lbvh.cpp:443-489 (32 in-place sweeps)
//   peach layout          core/builders/lbvh.cpp:84-188, 492-729, 805-824
//   BVH2 traversal        core/BVH2CommonRT.cpp:9-141 (+ raytrace_common.h:38-80)
//   any-hit rule          core/BVH4HalfRT.cpp:258-331 (tFar <= 0 => stop at first hit)
//   brute force           core/BruteForceRT.cpp:184-265
//   agglomerative tree    no reference counterpart (the engine's any-hit tree)
//   watertight test       no reference counterpart: Woop, Benthin and Wald,
//                         "Watertight Ray/Triangle Intersection", JCGT 2(1),
//                         2013 (section 3, the FP32 form without the
//                         double-precision recomputation of zero edges)
//   eye rays              examples/01_eye_rays/eye_ray.cpp:26-47, 114-138
//                         (kernel_InitEyeRay, SuperBlockIndex2DOpt,
//                         kernel_PackXY); examples/render_common.h:6-36
//                         (EyeRayDirNormalized, transform_ray3f).  LiteMath's
//                         float4x4 * float4 is taken as each row's products
//                         summed left to right, normalize(u) = u / sqrt(dot(u, u))
//   AO rays               examples/02_rtao/rtao.cpp:92, 147-165 (hit point,
//                         kernel_InitAORay); examples/render_common.h:56-112
//                         (MapSampleToCosineDistribution, OffsRayPos).  The
//                         draws are PCG32 (pcg-c-basic), sequence = global
//                         sample index, in place of the renderer's QMC table
//
// Floating point: every float3 operation is scalar, left to right
// (dot = x*x' + y*y' + z*z'), and the file is compiled with
// -ffp-contract=off and without fast-math, so each operation rounds once.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <unordered_set>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

static inline f3 sub3(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross3(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// std::min / std::max semantics (raytrace_common.h:76-79 uses them).
static inline float mn(float a, float b) { return (b < a) ? b : a; }
static inline float mx(float a, float b) { return (a < b) ? b : a; }

constexpr uint32_t START_MASK = 0x00FFFFFFu;   // raytrace_common.h:38
constexpr uint32_t SIZE_MASK = 0x7F000000u;    // raytrace_common.h:40
constexpr uint32_t LEAF_BIT = 0x80000000u;     // raytrace_common.h:41
constexpr float LBVH_INF = 1e10f;              // lbvh.cpp:53
constexpr float LBVH_EPS_BIAS = 1e-4f;         // lbvh.cpp:54
constexpr uint32_t PEACHES_MIN = 256;          // lbvh.h:25
constexpr uint32_t PEACHES_MAX = 2048;         // lbvh.h:26

static inline uint32_t pack_leaf(uint32_t start, uint32_t size) {  // raytrace_common.h:48-51
  return LEAF_BIT | ((size << 24) & SIZE_MASK) | (start & START_MASK);
}
static inline uint32_t leaf_start(uint32_t v) { return v & START_MASK; }
static inline uint32_t leaf_count(uint32_t v) { return (v & SIZE_MASK) >> 24; }

// 10-bit spread with the magic masks of lbvh.cpp:25-34.
static inline uint32_t spread_bits(int x, int offset) {
  x = (x | (x << 10)) & 0x000F801F;
  x = (x | (x << 4)) & 0x00E181C3;
  x = (x | (x << 2)) & 0x03248649;
  x = (x | (x << 2)) & 0x09249249;
  return uint32_t(x) << offset;
}
static inline uint32_t zindex3d(int x, int y, int z) {  // lbvh.cpp:36-39
  return spread_bits(x, 0) | spread_bits(y, 1) | spread_bits(z, 2);
}

// float -> int as the reference's implicit conversion at lbvh.cpp:282 does
// (truncation toward zero).  Values are in [0, 1024] for every input whose
// scene box is finite; out-of-range values are clamped the way the GPU's
// v_cvt_i32_f32 saturates, so both sides agree on pathological inputs.
static inline int f2i(float f) {
  if (!(f == f)) return 0;
  if (f >= 2147483647.0f) return 2147483647;
  if (f <= -2147483648.0f) return int(0x80000000u);
  return int(f);
}

struct Node {          // cbvh.h:20-26 (32 bytes)
  f3 bmin;
  uint32_t left;
  f3 bmax;
  uint32_t escape;
};

static inline Node dummy_node() {  // cbvh.h:28-36
  Node n;
  n.bmin = {0, 0, 0};
  n.bmax = {0, 0, 0};
  n.left = 0x80000000u;
  n.escape = 0xFFFFFFFFu;
  return n;
}

// BVH4 child (box + reference) and node, see collapse4 below
struct Child4 { f3 bmin, bmax; uint32_t ref; };
struct Quad { Child4 c[4]; };

struct Build {
  // inputs (float4 positions with w = 1, triangle indices)
  std::vector<f4> verts;
  std::vector<uint32_t> idx;
  uint32_t ntris = 0;
  // box mode (a top-level tree over instance boxes, BuildFromBoxes,
  // lbvh.cpp:829-917): 2 entries per primitive, min and max; verts/idx unused
  std::vector<f4> boxes;
  // cubic Morton grid: one scale for x, y and z, set by the largest extent
  // (the engine's any-hit tree); false = the reference's per-axis scale
  bool cubic = false;

  // builder state, named after lbvh.h:120-145
  f4 bbmin, bbmax, zscale;
  std::vector<uint32_t> codes_unsorted;  // code per triangle (input order)
  std::vector<uint32_t> sorted_codes;    // (code, tri) after the sort, first ntris entries
  std::vector<uint32_t> sorted_ids;
  std::vector<uint32_t> run_start;       // newIndices
  std::vector<uint32_t> ucodes;          // compressedCodes
  uint32_t L = 0;                        // leaves
  std::vector<Node> lr;                  // Karras left/right layout, 2L-1 nodes
  std::vector<uint32_t> perm;            // a_outIndices: leaf-order position -> triangle id
  std::vector<uint32_t> temp_count;      // m_tempCount
  uint32_t max_run = 0;

  // peach layout (BVH2_LEFT_OFFSET), as BVH2CommonRT traverses it
  std::vector<Node> peach;
  // leaf-order index buffer (AppendTreeData, BVH2CommonRT_host.cpp:45-64)
  std::vector<uint32_t> idx_leaf_order;
  // BVH4 collapse (FP32 boxes, FP16-rounded boxes), made on first use
  std::vector<Quad> q32, q16;
  // BVH8 collapse (FP16-rounded boxes), made on first use
};

static size_t power_of_two(size_t n) {  // lbvh.cpp:76-82
  size_t s = 2;
  while (s < n) s *= 2;
  return s;
}
static size_t round_up_to(size_t a, size_t n) {  // lbvh.cpp:69-74
  while (a % n != 0) a++;
  return a;
}

static int delta(const std::vector<uint32_t>& c, int i, int j) {  // lbvh.cpp:386-403
  if (j < 0 || j >= int(c.size())) return -1;
  uint32_t x = c[i] ^ c[j];
  if (x == 0) return 32;
  return __builtin_clz(x);
}

static inline int isign(int v) { return (v > 0) - (v < 0); }

static void karras(Build& b) {  // lbvh.cpp:405-441
  const std::vector<uint32_t>& c = b.ucodes;
  const int Lm1 = int(b.L) - 1;
  for (int i = 0; i < Lm1; ++i) {
    int d = isign(delta(c, i, i + 1) - delta(c, i, i - 1));
    int dmin = delta(c, i, i - d);
    int lmax = 2;
    while (delta(c, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t > 0; t >>= 1)
      if (delta(c, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(c, i, j);
    int s = 0;
    for (int t = (l + 1) >> 1; t > 0; t = (t == 1 ? 0 : (t + 1) >> 1))
      if (delta(c, i, i + (s + t) * d) > dnode) s += t;
    int gamma = i + s * d + std::min(d, 0);
    bool leftLeaf = (std::min(i, j) == gamma);
    bool rightLeaf = (std::max(i, j) == gamma + 1);
    b.lr[i].left = leftLeaf ? uint32_t(Lm1 + gamma) : uint32_t(gamma);
    b.lr[i].escape = rightLeaf ? uint32_t(Lm1 + gamma + 1) : uint32_t(gamma + 1);
  }
}

static void refit(Build& b) {  // lbvh.cpp:443-489, driver 781-785
  const int Lm1 = int(b.L) - 1;
  std::vector<uint32_t> lrx(Lm1 > 0 ? Lm1 : 0), lry(Lm1 > 0 ? Lm1 : 0);
  for (int i = 0; i < Lm1; i++) {
    lrx[i] = b.lr[i].left;
    lry[i] = b.lr[i].escape;
    b.temp_count[i] = 2;
  }
  int limit = int(b.L) - 1;  // int(newIndices.size()) - 1; set to -1 once the root is done
  for (int pass = 0; pass < 32; pass++) {
    for (int i = 0; i < Lm1; i++) {
      if (lrx[i] == 0xFFFFFFFFu) continue;
      uint32_t llx = 0xFFFFFFFFu, rlx = 0xFFFFFFFFu;
      if (int(lrx[i]) < limit) llx = lrx[lrx[i]];
      if (int(lry[i]) < limit) rlx = lrx[lry[i]];
      if (llx == 0xFFFFFFFFu && rlx == 0xFFFFFFFFu) {
        const Node& ln = b.lr[lrx[i]];
        const Node& rn = b.lr[lry[i]];
        Node cur;
        cur.bmin = {mn(ln.bmin.x, rn.bmin.x), mn(ln.bmin.y, rn.bmin.y), mn(ln.bmin.z, rn.bmin.z)};
        cur.bmax = {mx(ln.bmax.x, rn.bmax.x), mx(ln.bmax.y, rn.bmax.y), mx(ln.bmax.z, rn.bmax.z)};
        cur.left = lrx[i];
        cur.escape = lry[i];
        uint32_t lc = b.temp_count[lrx[i]], rc = b.temp_count[lry[i]];
        b.lr[i] = cur;
        lrx[i] = 0xFFFFFFFFu;
        lry[i] = 0xFFFFFFFFu;
        b.temp_count[i] = lc + rc + 2;
        if (i == 0) limit = -1;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Peach layout (lbvh.cpp:84-188 Reserve/BuildPeachTree, 492-729 conversion)
// ---------------------------------------------------------------------------
struct U2 { uint32_t x, y; };

static void push_down_escape(std::vector<U2>& le, uint32_t off, uint32_t parentEsc) {  // lbvh.cpp:141-154
  if (off >= le.size()) return;
  const U2 l = le[off + 0];
  const U2 r = le[off + 1];
  le[off + 0].y = off + 1;
  le[off + 1].y = parentEsc;
  push_down_escape(le, l.x, off + 1);
  push_down_escape(le, r.x, parentEsc);
}

static void build_peach(Build& b) {
  const size_t nprims = b.ntris;
  const size_t childs = 2;
  size_t currOffs = 0, currSize = 2;
  size_t peachSize = (nprims <= PEACHES_MAX * 4) ? PEACHES_MIN : PEACHES_MAX;
  std::vector<U2> intervals;
  std::vector<uint32_t> countPrefix;
  if (nprims == 1) {
    intervals.push_back({0, uint32_t(childs)});
    currOffs = childs;
  } else {
    do {
      intervals.push_back({uint32_t(currOffs), uint32_t(currSize)});
      currOffs += currSize;
      currSize *= childs;
    } while (currSize < peachSize);
    countPrefix.assign(currSize, 0);
    intervals.push_back({uint32_t(currOffs), uint32_t(currSize)});
    currOffs += currSize;
    currSize *= childs;
    currOffs += currSize;
    currOffs += round_up_to(2 * nprims - 1, childs);
  }
  const size_t reserved = currOffs;
  std::vector<U2> le(reserved, U2{0xFFFFFFFFu, 0});
  for (size_t lev = 0; lev + 1 < intervals.size(); lev++) {
    const U2 cur = intervals[lev], nxt = intervals[lev + 1];
    for (uint32_t i = cur.x; i < cur.x + cur.y; i++) le[i] = U2{nxt.x + 2 * (i - cur.x), 0};
  }
  push_down_escape(le, 0, uint32_t(-2));

  std::vector<U2> addr(reserved, U2{0, 0});
  std::vector<Node> out(reserved);
  for (auto& n : out) { n.bmin = {0, 0, 0}; n.bmax = {0, 0, 0}; n.left = 0; n.escape = 0; }
  const std::vector<Node>& in = b.lr;

  // PeachConvertInit, lbvh.cpp:492-543
  {
    Node root = in[0];
    if (int(root.left) < 0) {
      root.escape = uint32_t(-2);
      addr[0] = U2{0xFFFFFFFFu, 0xFFFFFFFFu};
      addr[1] = U2{0xFFFFFFFFu, 0xFFFFFFFFu};
      out[0] = root;
      out[1] = dummy_node();
    } else {
      const Node left = in[root.left];
      const Node right = in[root.escape];
      addr[0] = U2{left.left, left.escape};
      addr[1] = U2{right.left, right.escape};
      Node ln = out[0], rn = out[1];
      ln.bmin = left.bmin; ln.bmax = left.bmax;
      rn.bmin = right.bmin; rn.bmax = right.bmax;
      ln.left = le[0].x; ln.escape = le[0].y;
      rn.left = le[1].x; rn.escape = le[1].y;
      if (int(left.left) < 0) { ln.left = left.left; addr[0] = U2{0xFFFFFFFFu, 0xFFFFFFFFu}; }
      if (int(right.left) < 0) { rn.left = right.left; addr[1] = U2{0xFFFFFFFFu, 0xFFFFFFFFu}; }
      out[0] = ln;
      out[1] = rn;
    }
  }
  // PeachConvertPass, lbvh.cpp:545-612
  for (size_t lev = 0; lev + 1 < intervals.size(); lev++) {
    const U2 cur = intervals[lev], nxt = intervals[lev + 1];
    const bool last = (lev == intervals.size() - 2);
    for (uint32_t i = 0; i < cur.y; i++) {
      const U2 lr = addr[cur.x + i];
      bool leftLeaf = true, rightLeaf = true;
      if (lr.x == 0xFFFFFFFFu) {
        out[nxt.x + 2 * i + 0] = dummy_node();
        addr[nxt.x + 2 * i + 0] = U2{0xFFFFFFFFu, 0xFFFFFFFFu};
      } else {
        const Node left = in[lr.x];
        const U2 lle = le[nxt.x + 2 * i + 0];
        leftLeaf = (int(left.left) < 0);
        Node n;
        n.bmin = left.bmin; n.bmax = left.bmax;
        n.left = leftLeaf ? left.left : lle.x;
        n.escape = lle.y;
        out[nxt.x + 2 * i + 0] = n;
        addr[nxt.x + 2 * i + 0] = leftLeaf ? U2{0xFFFFFFFFu, 0xFFFFFFFFu} : U2{left.left, left.escape};
      }
      if (lr.y == 0xFFFFFFFFu) {
        out[nxt.x + 2 * i + 1] = dummy_node();
        addr[nxt.x + 2 * i + 1] = U2{0xFFFFFFFFu, 0xFFFFFFFFu};
      } else {
        const Node right = in[lr.y];
        const U2 rle = le[nxt.x + 2 * i + 1];
        rightLeaf = (int(right.left) < 0);
        Node n;
        n.bmin = right.bmin; n.bmax = right.bmax;
        n.left = rightLeaf ? right.left : rle.x;
        n.escape = rle.y;
        out[nxt.x + 2 * i + 1] = n;
        addr[nxt.x + 2 * i + 1] = rightLeaf ? U2{0xFFFFFFFFu, 0xFFFFFFFFu} : U2{right.left, right.escape};
      }
      if (last) {
        countPrefix[2 * i + 0] = leftLeaf ? 0 : b.temp_count[lr.x];
        countPrefix[2 * i + 1] = rightLeaf ? 0 : b.temp_count[lr.y];
      }
    }
  }
  for (size_t i = 1; i < countPrefix.size(); i++) countPrefix[i] += countPrefix[i - 1];  // inclusive scan, lbvh.cpp:816

  // RecursiveConvert, lbvh.cpp:644-729
  if (intervals.size() != 1) {
    const U2 lastLev = intervals.back();
    for (uint32_t i = 0; i < lastLev.y; i++) {
      const U2 lr = addr[lastLev.x + i];
      uint32_t sub = lastLev.x + lastLev.y;
      if (i > 0) sub += countPrefix[i - 1];
      if (lr.x == 0xFFFFFFFFu) continue;
      out[lastLev.x + i].left = sub;
      struct U4 { uint32_t x, y, z, w; } stack[32];
      int top = 0;
      uint32_t lo = lr.x, ro = lr.y, pesc = out[lastLev.x + i].escape;
      while (top >= 0) {
        bool pop = (int(lo) < 0);
        if (int(lo) > 0) {
          const Node ln = in[lo];
          const Node rn = in[ro];
          const uint32_t cs = sub;
          sub += 2;
          out[cs + 0] = ln;
          out[cs + 1] = rn;
          out[cs + 0].escape = cs + 1;
          out[cs + 1].escape = pesc;
          pop = (int(ln.left) < 0) && (int(rn.left) < 0);
          if (int(ln.left) > 0 && int(rn.left) > 0) {
            stack[top] = U4{rn.left, rn.escape, pesc, cs + 1};
            top++;
          }
          if (int(ln.left) > 0) {
            lo = ln.left; ro = ln.escape; pesc = cs + 1;
            out[cs + 0].left = sub;
          } else if (int(rn.left) > 0) {
            lo = rn.left; ro = rn.escape;
            out[cs + 1].left = sub;
          }
        }
        if (pop) {
          top--;
          U4 od = stack[std::max(top, 0)];
          lo = od.x; ro = od.y; pesc = od.z;
          if (int(lo) > 0 && top >= 0) out[od.w].left = sub;
        }
      }
    }
  }
  if (out.size() % 2 != 0) out.push_back(dummy_node());  // cbvh_hlbvh.cpp:232-236
  b.peach.swap(out);
}

static void build_all(Build& b) {
  const uint32_t N = b.ntris;
  // scene box over all vertices, lbvh.cpp:228-239
  const bool box_mode = !b.boxes.empty();
  f4 lo = {LBVH_INF, LBVH_INF, LBVH_INF, LBVH_INF}, hi = {-LBVH_INF, -LBVH_INF, -LBVH_INF, -LBVH_INF};
  if (box_mode) {  // over the primitive boxes, lbvh.cpp:190-203
    for (uint32_t i = 0; i < N; i++) {
      const f4 a = b.boxes[2 * i], c = b.boxes[2 * i + 1];
      lo = {mn(lo.x, a.x), mn(lo.y, a.y), mn(lo.z, a.z), mn(lo.w, a.w)};
      hi = {mx(hi.x, c.x), mx(hi.y, c.y), mx(hi.z, c.z), mx(hi.w, c.w)};
    }
  } else {
    for (const f4& v : b.verts) {
      lo = {mn(lo.x, v.x), mn(lo.y, v.y), mn(lo.z, v.z), mn(lo.w, v.w)};
      hi = {mx(hi.x, v.x), mx(hi.y, v.y), mx(hi.z, v.z), mx(hi.w, v.w)};
    }
  }
  b.bbmin = lo;
  b.bbmax = hi;
  // ZIndexMultCoeff(bboxMin - 1e-4, bboxMax + 1e-4), lbvh.cpp:41-51, 263
  {
    const float ex[4] = {(hi.x + LBVH_EPS_BIAS) - (lo.x - LBVH_EPS_BIAS), (hi.y + LBVH_EPS_BIAS) - (lo.y - LBVH_EPS_BIAS),
                         (hi.z + LBVH_EPS_BIAS) - (lo.z - LBVH_EPS_BIAS), (hi.w + LBVH_EPS_BIAS) - (lo.w - LBVH_EPS_BIAS)};
    float inv[4];
    for (int k = 0; k < 4; k++) inv[k] = 1.0f / ex[k];
    for (int k = 0; k < 3; k++)
      if (std::fabs(ex[k]) < 1e-20f) inv[k] = 0.0f;
    if (b.cubic) {
      const float m = mx(ex[0], mx(ex[1], ex[2]));
      const float im = (std::fabs(m) < 1e-20f) ? 0.0f : 1.0f / m;
      inv[0] = inv[1] = inv[2] = im;
    }
    b.zscale = {1024.0f * inv[0], 1024.0f * inv[1], 1024.0f * inv[2], 1024.0f * inv[3]};
  }
  // Morton codes, lbvh.cpp:259-287
  b.codes_unsorted.resize(N);
  for (uint32_t i = 0; i < N; i++) {
    float cx, cy, cz;
    if (box_mode) {  // box centre, lbvh.cpp:251
      const f4 a = b.boxes[2 * i], c = b.boxes[2 * i + 1];
      cx = (a.x + c.x) * 0.5f;
      cy = (a.y + c.y) * 0.5f;
      cz = (a.z + c.z) * 0.5f;
      const float sx = b.zscale.x * (cx - lo.x);
      const float sy = b.zscale.y * (cy - lo.y);
      const float sz = b.zscale.z * (cz - lo.z);
      b.codes_unsorted[i] = zindex3d(f2i(sx), f2i(sy), f2i(sz));
      continue;
    }
    const f4 v1 = b.verts[b.idx[3 * i + 0]], v2 = b.verts[b.idx[3 * i + 1]], v3 = b.verts[b.idx[3 * i + 2]];
    if (N < 256) {
      cx = 0.33333333f * ((v1.x + v2.x) + v3.x);
      cy = 0.33333333f * ((v1.y + v2.y) + v3.y);
      cz = 0.33333333f * ((v1.z + v2.z) + v3.z);
    } else {
      cx = (mn(v1.x, mn(v2.x, v3.x)) + mx(v1.x, mx(v2.x, v3.x))) * 0.5f;
      cy = (mn(v1.y, mn(v2.y, v3.y)) + mx(v1.y, mx(v2.y, v3.y))) * 0.5f;
      cz = (mn(v1.z, mn(v2.z, v3.z)) + mx(v1.z, mx(v2.z, v3.z))) * 0.5f;
    }
    const float sx = b.zscale.x * (cx - lo.x);
    const float sy = b.zscale.y * (cy - lo.y);
    const float sz = b.zscale.z * (cz - lo.z);
    b.codes_unsorted[i] = zindex3d(f2i(sx), f2i(sy), f2i(sz));
  }
  // sort by (code, index): stable
  std::vector<uint32_t> order(N);
  for (uint32_t i = 0; i < N; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t c) { return b.codes_unsorted[a] < b.codes_unsorted[c]; });
  b.sorted_codes.resize(N);
  b.sorted_ids.resize(N);
  for (uint32_t i = 0; i < N; i++) {
    b.sorted_ids[i] = order[i];
    b.sorted_codes[i] = b.codes_unsorted[order[i]];
  }
  // unique-code runs, lbvh.cpp:289-319
  b.run_start.clear();
  b.ucodes.clear();
  for (uint32_t i = 0; i < N; i++) {
    if (i == 0 || b.sorted_codes[i - 1] != b.sorted_codes[i]) {
      b.run_start.push_back(i);
      b.ucodes.push_back(b.sorted_codes[i]);
    }
  }
  b.L = uint32_t(b.run_start.size());
  const uint32_t L = b.L;
  b.lr.assign(2 * size_t(L) - 1, Node{{0, 0, 0}, 0, {0, 0, 0}, 0});
  b.perm.assign(N, 0);
  b.temp_count.assign(2 * size_t(N) - 1, 0);
  b.max_run = 0;
  // leaves, lbvh.cpp:350-384
  for (uint32_t i = 0; i < L; i++) {
    const uint32_t start = b.run_start[i];
    const uint32_t count = (i < L - 1) ? b.run_start[i + 1] - start : N - b.run_start[L - 1];
    b.max_run = std::max(b.max_run, count);
    f4 bl = {LBVH_INF, LBVH_INF, LBVH_INF, LBVH_INF}, bh = {-LBVH_INF, -LBVH_INF, -LBVH_INF, -LBVH_INF};
    for (uint32_t j = start; j < start + count; j++) {
      const uint32_t t = b.sorted_ids[j];
      b.perm[j] = t;
      if (box_mode) {  // union of the primitive boxes (lbvh.cpp:321-348, every primitive of the run kept)
        const f4 a = b.boxes[2 * t], c = b.boxes[2 * t + 1];
        bl = {mn(bl.x, a.x), mn(bl.y, a.y), mn(bl.z, a.z), 0};
        bh = {mx(bh.x, c.x), mx(bh.y, c.y), mx(bh.z, c.z), 0};
        continue;
      }
      const f4 v1 = b.verts[b.idx[3 * t + 0]], v2 = b.verts[b.idx[3 * t + 1]], v3 = b.verts[b.idx[3 * t + 2]];
      bl = {mn(bl.x, mn(v1.x, mn(v2.x, v3.x))), mn(bl.y, mn(v1.y, mn(v2.y, v3.y))), mn(bl.z, mn(v1.z, mn(v2.z, v3.z))), 0};
      bh = {mx(bh.x, mx(v1.x, mx(v2.x, v3.x))), mx(bh.y, mx(v1.y, mx(v2.y, v3.y))), mx(bh.z, mx(v1.z, mx(v2.z, v3.z))), 0};
    }
    Node& n = b.lr[L - 1 + i];
    n.bmin = {bl.x, bl.y, bl.z};
    n.bmax = {bh.x, bh.y, bh.z};
    n.left = pack_leaf(start, count);
    n.escape = 0xFFFFFFFFu;
  }
  karras(b);
  refit(b);
  if (box_mode) return;
  build_peach(b);
  // AppendTreeData: index buffer rewritten in leaf order (BVH2CommonRT_host.cpp:54-61)
  b.idx_leaf_order.resize(3 * size_t(N));
  for (uint32_t i = 0; i < N; i++)
    for (int k = 0; k < 3; k++) b.idx_leaf_order[3 * i + k] = b.idx[3 * b.perm[i] + k];
}

// ---------------------------------------------------------------------------
// Traversal
// ---------------------------------------------------------------------------
struct Hit {  // CRT_Hit (CrossRT.h:20-27), primId/instId/geomId as u32
  float t;
  uint32_t primId, instId, geomId;
  float u, v;
};

static inline f3 safe_inverse(f3 d) {  // raytrace_common.h:57-65
  const float eps = 1.0e-36f;
  return {1.0f / (std::fabs(d.x) > eps ? d.x : std::copysign(eps, d.x)),
          1.0f / (std::fabs(d.y) > eps ? d.y : std::copysign(eps, d.y)),
          1.0f / (std::fabs(d.z) > eps ? d.z : std::copysign(eps, d.z))};
}

// The triangle test every traversal below uses: 0 = the reference's
// Moller-Trumbore, 1 = the watertight test (which brings the conservative
// box test slab_robust with it).  Set by the Python wrapper for the duration
// of one call (hrto_set_triangle_test); not for concurrent calls that want
// different tests.
static int g_tri_test = 0;

static inline void slab_ref(f3 o, f3 inv, f3 bmin, f3 bmax, float& tmin_out, float& tmax_out) {  // raytrace_common.h:67-80
  const float lo = inv.x * (bmin.x - o.x);
  const float hi = inv.x * (bmax.x - o.x);
  const float lo1 = inv.y * (bmin.y - o.y);
  const float hi1 = inv.y * (bmax.y - o.y);
  const float lo2 = inv.z * (bmin.z - o.z);
  const float hi2 = inv.z * (bmax.z - o.z);
  const float tmin = mx(mn(lo, hi), mn(lo1, hi1));
  const float tmax = mn(mx(lo, hi), mx(lo1, hi1));
  tmin_out = mx(tmin, mn(lo2, hi2));
  tmax_out = mn(tmax, mx(lo2, hi2));
}

// The engine's fused form of the same test for the any-hit traversals of FP16
// nodes: t = fma(b, inv, -(o * inv)) per plane (one rounding after the
// product o * inv).  Per axis the bound the ray meets first is the minimum
// where inv >= 0, else the maximum; entry = the largest of those distances,
// exit = the smallest of the others.  For a box with min <= max this equals
// the min/max reduction (rounding is monotonic); an inverted box (the
// +65504 .. -65504 of an unused slot) comes out empty.  Used only for a ray
// whose -(o * inv) is finite on every axis (fused_ok); other rays keep slab.
static inline void slab_fused(f3 o, f3 inv, f3 bmin, f3 bmax, float& tmin_out, float& tmax_out) {
  const float nx = -(o.x * inv.x), ny = -(o.y * inv.y), nz = -(o.z * inv.z);
  const float ax = std::fma(inv.x >= 0.0f ? bmin.x : bmax.x, inv.x, nx);
  const float bx = std::fma(inv.x >= 0.0f ? bmax.x : bmin.x, inv.x, nx);
  const float ay = std::fma(inv.y >= 0.0f ? bmin.y : bmax.y, inv.y, ny);
  const float by = std::fma(inv.y >= 0.0f ? bmax.y : bmin.y, inv.y, ny);
  const float az = std::fma(inv.z >= 0.0f ? bmin.z : bmax.z, inv.z, nz);
  const float bz = std::fma(inv.z >= 0.0f ? bmax.z : bmin.z, inv.z, nz);
  tmin_out = mx(mx(ax, ay), az);
  tmax_out = mn(mn(bx, by), bz);
}

// The box test of the watertight mode, conservative as in Ize, "Robust BVH
// Ray Traversal" (JCGT 2(2), 2013): each plane distance inv * (b - o) has three
// roundings (reciprocal, difference, product), so its relative error is within
// gamma3 = 3u / (1 - 3u), u = 2^-24; the exit distance is multiplied by
// 1 + 2 gamma3, taken as the next float above (1 + 4 * 2^-23).  An axis whose
// direction component is within +-1e-36 (where safe_inverse stands in +-1e36
// for 1/d) counts as parallel: all t when lo <= o <= hi, none otherwise.
static const float kRobustExit = 1.000000476837158203125f;
static inline void robust_axis(float o, float inv, float d, float lo, float hi, float& a, float& b) {
  const float t0 = inv * (lo - o), t1 = inv * (hi - o);
  a = mn(t0, t1);
  b = mx(t0, t1);
  if (!(std::fabs(d) > 1.0e-36f)) {
    const bool in = lo <= o && o <= hi;
    a = in ? -INFINITY : INFINITY;
    b = in ? INFINITY : -INFINITY;
  }
}
static inline void slab_robust(f3 o, f3 d, f3 inv, f3 bmin, f3 bmax, float& tmin_out, float& tmax_out) {
  float ax, bx, ay, by, az, bz;
  robust_axis(o.x, inv.x, d.x, bmin.x, bmax.x, ax, bx);
  robust_axis(o.y, inv.y, d.y, bmin.y, bmax.y, ay, by);
  robust_axis(o.z, inv.z, d.z, bmin.z, bmax.z, az, bz);
  tmin_out = mx(mx(ax, ay), az);
  tmax_out = mn(mn(bx, by), bz) * kRobustExit;
}

// The box test of the current triangle test.
static inline void slab(f3 o, f3 d, f3 inv, f3 bmin, f3 bmax, float& tmin_out, float& tmax_out) {
  if (g_tri_test == 1) slab_robust(o, d, inv, bmin, bmax, tmin_out, tmax_out);
  else slab_ref(o, inv, bmin, bmax, tmin_out, tmax_out);
}

static inline bool fused_ok(f3 o, f3 inv) {
  return std::isfinite(-(o.x * inv.x)) && std::isfinite(-(o.y * inv.y)) && std::isfinite(-(o.z * inv.z));
}

// The watertight test (Woop, Benthin, Wald 2013) in FP32, the published
// steps: the ray's largest |direction| axis becomes z (the lowest axis on
// ties; x and y exchanged when that component is negative), the vertices are
// taken relative to the origin and sheared by (-dx/dz, -dy/dz, 1/dz), the 2D
// edge functions U (edge BC), V (CA), W (AB) must not have mixed signs (a
// zero counts as either), det = U + V + W must not be zero, and
// t = (U Az + V Bz + W Cz) / det.  u = W / det (the weight of C) and
// v = V / det (the weight of B), the reference's meaning of u and v.  Every
// operation rounds once, in the order written.
static inline float axis_of(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

static inline bool wt_eval(f3 o, f3 d, float tNear, f3 A0, f3 B0, f3 C0, float& t, float& u, float& v) {
  const float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
  const int kz = (ax >= ay) ? (ax >= az ? 0 : 2) : (ay >= az ? 1 : 2);
  int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
  const float dz = axis_of(d, kz);
  if (dz < 0.0f) std::swap(kx, ky);
  const float sx = axis_of(d, kx) / dz, sy = axis_of(d, ky) / dz, sz = 1.0f / dz;
  const f3 A = sub3(A0, o), B = sub3(B0, o), C = sub3(C0, o);
  const float Ax = axis_of(A, kx) - sx * axis_of(A, kz), Ay = axis_of(A, ky) - sy * axis_of(A, kz);
  const float Bx = axis_of(B, kx) - sx * axis_of(B, kz), By = axis_of(B, ky) - sy * axis_of(B, kz);
  const float Cx = axis_of(C, kx) - sx * axis_of(C, kz), Cy = axis_of(C, ky) - sy * axis_of(C, kz);
  const float U = Cx * By - Cy * Bx;
  const float V = Ax * Cy - Ay * Cx;
  const float W = Bx * Ay - By * Ax;
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = U + V + W;
  if (det == 0.0f) return false;
  const float T = U * (sz * axis_of(A, kz)) + V * (sz * axis_of(B, kz)) + W * (sz * axis_of(C, kz));
  const float rcp = 1.0f / det;
  t = T * rcp;
  u = W * rcp;
  v = V * rcp;
  return t > tNear;
}

// Moller-Trumbore with the reference's 1e-6 slack, BVH2CommonRT.cpp:26-45:
// t, u, v and whether the barycentric and t > tNear conditions hold.
static inline bool mt_eval(f3 o, f3 d, float tNear, f3 A, f3 B, f3 C, float& t, float& u, float& v) {
  if (g_tri_test == 1) return wt_eval(o, d, tNear, A, B, C, t, u, v);
  const f3 e1 = sub3(B, A);
  const f3 e2 = sub3(C, A);
  const f3 pv = cross3(d, e2);
  const f3 tv = sub3(o, A);
  const f3 qv = cross3(tv, e1);
  const float invDet = 1.0f / dot3(e1, pv);
  v = dot3(tv, pv) * invDet;
  u = dot3(qv, d) * invDet;
  t = dot3(e2, qv) * invDet;
  return v >= -1e-6f && u >= -1e-6f && (u + v <= 1.0f + 1e-6f) && t > tNear;
}

// The reference's acceptance: first strictly closer hit (BVH2CommonRT.cpp:37).
// Returns true and updates the hit when accepted.
static inline bool mt_test(f3 o, f3 d, float tNear, f3 A, f3 B, f3 C, Hit& h, uint32_t prim) {
  float t, u, v;
  if (mt_eval(o, d, tNear, A, B, C, t, u, v) && t < h.t) {
    h.t = t;
    h.primId = prim;
    h.instId = 0;
    h.geomId = 0;
    h.u = u;
    h.v = v;
    return true;
  }
  return false;
}

static inline f3 vpos(const Build& b, uint32_t vi) {
  const f4 v = b.verts[vi];
  return {v.x, v.y, v.z};
}

struct Counters { uint64_t pairs = 0, tris = 0, leaves = 0, max_stack = 0; };

// The reference's traversal metrics as shipped: CrossRT.h:31 defines
// ENABLE_METRICS, and BVH2TraverseC32 then keeps, per visited node pair,
// NC/BLB and -- for each of the five treelet sizes (CrossRT.h:32-33) -- long
// jumps, cache-line changes and a set of touched lines (BVH2CommonRT.cpp:
// 57-84), per leaf LC/TC/BLB (:121-126), per push and pop SOC/SBL
// (:108-111, :134-138).  Only the CPU baseline's "as shipped" timing uses it.
struct RefMetrics {
  static constexpr int kTreelets = 5;
  static constexpr uint32_t kSizes[kTreelets] = {128, 512, 4096, 16384, 131072};
  uint64_t NC = 0, LC = 0, TC = 0, BLB = 0, SOC = 0, SBL = 0;
  uint64_t LJC[kTreelets] = {0}, CMC[kTreelets] = {0};
  std::unordered_set<uint32_t> WSS[kTreelets];
};
constexpr uint32_t RefMetrics::kSizes[RefMetrics::kTreelets];

// The traversal loop of BVH2CommonRT::BVH2TraverseC32 in object space, with
// an incoming hit (culling uses its distance) and a triangle acceptance rule
// accept(A, B, C, original id) that updates the hit.  Returns true when an
// any-hit query has found its hit.
template <class Accept>
static bool visit_bvh2(const Build& b, f3 o, f3 d, float tNear, bool anyhit, Hit& h, Counters* cnt,
                       Accept&& accept, RefMetrics* m = nullptr) {
  const f3 inv = safe_inverse(d);
  uint32_t stack[80];
  int top = 0;
  uint32_t off = 0;
  uint32_t off_old = 0;
  const Node* nodes = b.peach.data();
  while (top >= 0) {
    bool pop = false;
    if ((off & LEAF_BIT) == 0) {
      if (m) {
        m->NC += 2;
        m->BLB += 2 * sizeof(Node);
        if (off_old != off) {
          for (int i = 0; i < RefMetrics::kTreelets; i++) {
            const double jump = std::abs(double(off) - double(off_old)) * double(sizeof(Node));
            if (jump >= double(RefMetrics::kSizes[i])) m->LJC[i]++;
            const uint32_t line_old = uint32_t(off_old * sizeof(Node)) / RefMetrics::kSizes[i];
            const uint32_t line_new = uint32_t(off * sizeof(Node)) / RefMetrics::kSizes[i];
            if (line_old != line_new) {
              m->CMC[i]++;
              m->WSS[i].insert(line_new);
            }
          }
          off_old = off;
        }
      }
      const Node n0 = nodes[off + 0];
      const Node n1 = nodes[off + 1];
      if (cnt) cnt->pairs++;
      float a0, b0, a1, b1;
      slab(o, d, inv, n0.bmin, n0.bmax, a0, b0);
      slab(o, d, inv, n1.bmin, n1.bmax, a1, b1);
      const bool h0 = (a0 <= b0) && (b0 >= tNear) && (a0 <= h.t);
      const bool h1 = (a1 <= b1) && (b1 >= tNear) && (a1 <= h.t);
      pop = (!h0 && !h1);
      off = h0 ? n0.left : n1.left;
      if (h0 && h1) {
        off = (a0 <= a1) ? n0.left : n1.left;
        stack[top] = (a0 <= a1) ? n1.left : n0.left;
        top++;
        if (cnt && uint64_t(top) > cnt->max_stack) cnt->max_stack = uint64_t(top);
        if (m) {
          m->SOC++;
          m->SBL += sizeof(uint32_t);
        }
      }
    } else if (off != 0xFFFFFFFFu) {
      const uint32_t start = leaf_start(off), count = leaf_count(off);
      if (cnt) { cnt->leaves++; cnt->tris += count; }
      if (m) {
        m->LC++;
        m->TC += count;
        m->BLB += count * (3 * sizeof(uint32_t) + 3 * sizeof(f4));
      }
      for (uint32_t t = start; t < start + count; t++) {
        const uint32_t* ix = &b.idx_leaf_order[3 * size_t(t)];
        accept(vpos(b, ix[0]), vpos(b, ix[1]), vpos(b, ix[2]), b.perm[t]);
      }
      pop = true;
      if (anyhit && h.primId != 0xFFFFFFFFu) return true;
    }
    if (pop) {
      top--;
      off = stack[std::max(top, 0)];
      if (m) {
        m->SOC++;
        m->SBL += sizeof(uint32_t);
      }
    }
  }
  return false;
}

// BVH2CommonRT::BVH2TraverseC32 over the peach layout (BVH2CommonRT.cpp:49-141).
// primId is reported as the ORIGINAL triangle id (the REMAP_PRIM_ID rule of
// BVH2CommonRT.cpp:270-276); the leaf-order position is perm^-1 of it.
static Hit trace_bvh2common(const Build& b, f4 posNear, f4 dirFar, Counters* cnt, RefMetrics* m = nullptr) {
  bool anyhit = (dirFar.w <= 0.0f);
  float tFar = anyhit ? -dirFar.w : dirFar.w;
  Hit h;
  h.t = tFar;
  h.primId = h.instId = h.geomId = 0xFFFFFFFFu;
  h.u = h.v = 0.0f;
  const f3 o = {posNear.x, posNear.y, posNear.z};
  const f3 d = {dirFar.x, dirFar.y, dirFar.z};
  const float tNear = posNear.w;
  visit_bvh2(b, o, d, tNear, anyhit, h, cnt,
             [&](f3 A, f3 B, f3 C, uint32_t prim) { mt_test(o, d, tNear, A, B, C, h, prim); }, m);
  return h;
}

// BVH2_LRFT-style traversal directly over the Karras left/right layout
// (BVH2CommonLRFT.cpp:81-221 with a single BLAS and no TLAS).  Visits the
// same (left, right) children in the same near-first order as the peach
// traversal, so it must give bit-identical hits; the tests use that to show
// the hits do not depend on the node layout.
static Hit trace_lr(const Build& b, f4 posNear, f4 dirFar) {
  bool anyhit = (dirFar.w <= 0.0f);
  float tFar = anyhit ? -dirFar.w : dirFar.w;
  Hit h;
  h.t = tFar;
  h.primId = h.instId = h.geomId = 0xFFFFFFFFu;
  h.u = h.v = 0.0f;
  const f3 o = {posNear.x, posNear.y, posNear.z};
  const f3 d = {dirFar.x, dirFar.y, dirFar.z};
  const float tNear = posNear.w;
  const f3 inv = safe_inverse(d);
  const Node* nodes = b.lr.data();
  if (b.L == 1) {  // root is a leaf: its box is tested first, as in the peach layout
    const uint32_t leaf = nodes[0].left;
    float a0, b0;
    slab(o, d, inv, nodes[0].bmin, nodes[0].bmax, a0, b0);
    if (!((a0 <= b0) && (b0 >= tNear) && (a0 <= h.t))) return h;
    for (uint32_t t = leaf_start(leaf); t < leaf_start(leaf) + leaf_count(leaf); t++) {
      const uint32_t* ix = &b.idx_leaf_order[3 * size_t(t)];
      mt_test(o, d, tNear, vpos(b, ix[0]), vpos(b, ix[1]), vpos(b, ix[2]), h, b.perm[t]);
    }
    return h;
  }
  uint32_t stack[80];
  int top = 0;
  uint32_t cur = 0;  // internal node whose two children are tested next
  while (true) {
    const Node& p = nodes[cur];
    const Node n0 = nodes[p.left];
    const Node n1 = nodes[p.escape];
    float a0, b0, a1, b1;
    slab(o, d, inv, n0.bmin, n0.bmax, a0, b0);
    slab(o, d, inv, n1.bmin, n1.bmax, a1, b1);
    const bool h0 = (a0 <= b0) && (b0 >= tNear) && (a0 <= h.t);
    const bool h1 = (a1 <= b1) && (b1 >= tNear) && (a1 <= h.t);
    uint32_t next = 0xFFFFFFFFu;
    if (h0 && h1) {
      next = (a0 <= a1) ? p.left : p.escape;
      stack[top++] = (a0 <= a1) ? p.escape : p.left;
    } else if (h0) {
      next = p.left;
    } else if (h1) {
      next = p.escape;
    }
    // descend until an internal node is found or the stack runs out
    while (true) {
      if (next == 0xFFFFFFFFu) {
        if (top == 0) return h;
        next = stack[--top];
      }
      const uint32_t lo = nodes[next].left;
      if (lo & LEAF_BIT) {
        for (uint32_t t = leaf_start(lo); t < leaf_start(lo) + leaf_count(lo); t++) {
          const uint32_t* ix = &b.idx_leaf_order[3 * size_t(t)];
          mt_test(o, d, tNear, vpos(b, ix[0]), vpos(b, ix[1]), vpos(b, ix[2]), h, b.perm[t]);
        }
        if (anyhit && h.primId != 0xFFFFFFFFu) return h;
        next = 0xFFFFFFFFu;
        continue;
      }
      cur = next;
      break;
    }
  }
}

// BruteForceRT::RayQuery_NearestHit for one geometry and one identity
// instance (BruteForceRT.cpp:184-265): every triangle in input order, strict
// t < hit.t, original triangle id.
static Hit trace_brute(const Build& b, f4 posNear, f4 dirFar) {
  bool anyhit = (dirFar.w <= 0.0f);
  float tFar = anyhit ? -dirFar.w : dirFar.w;
  Hit h;
  h.t = tFar;
  h.primId = h.instId = h.geomId = 0xFFFFFFFFu;
  h.u = h.v = 0.0f;
  const f3 o = {posNear.x, posNear.y, posNear.z};
  const f3 d = {dirFar.x, dirFar.y, dirFar.z};
  const float tNear = posNear.w;
  // instance box test (BruteForceRT.cpp:196-201) against the geometry box
  {
    f3 lo = {b.bbmin.x, b.bbmin.y, b.bbmin.z}, hi = {b.bbmax.x, b.bbmax.y, b.bbmax.z};
    float a, c;
    slab(o, d, safe_inverse(d), lo, hi, a, c);
    if (!((a <= c) && (c >= tNear) && (a <= tFar))) return h;
  }
  for (uint32_t t = 0; t < b.ntris; t++) {
    mt_test(o, d, tNear, vpos(b, b.idx[3 * t]), vpos(b, b.idx[3 * t + 1]), vpos(b, b.idx[3 * t + 2]), h, t);
    if (anyhit && h.primId != 0xFFFFFFFFu) break;
  }
  return h;
}

// ---------------------------------------------------------------------------
// Agglomerative hierarchy over the leaves (PLOC: Meister & Bittner 2018,
// "Parallel Locally-Ordered Clustering for Bounding Volume Hierarchy
// Construction").  No reference counterpart: the engine's any-hit tree may
// differ from the reference's tree, since an any-hit answer does not depend
// on the tree.  Restated so that the GPU builder's nodes can be checked one
// for one.  The rule, step by step:
//   clusters = the leaves in Morton order, each with its box;
//   repeat until one cluster is left:
//     each cluster i picks the j != i within [i - radius, i + radius] that
//     minimises the union box's half area (ex*ey + ey*ez + ez*ex, NaN taken
//     as +inf), ties to the smaller j -- a strict order on pairs, so the
//     globally best pair always picks itself and every round merges;
//     every mutual pair (i < j) becomes a new internal node, left = i,
//     right = j, numbered downwards from L - 2 in increasing i; it takes
//     position i and j is dropped; the survivors keep their order.
// The root ends up as node 0.  Boxes are the exact min/max unions, so the
// LR layout (internal nodes 0..L-2, leaves L-1+m unchanged) feeds the BVH4
// collapse as a Karras tree does.
// ---------------------------------------------------------------------------
static inline float union_half_area(f3 alo, f3 ahi, f3 blo, f3 bhi) {
  const f3 lo = {mn(alo.x, blo.x), mn(alo.y, blo.y), mn(alo.z, blo.z)};
  const f3 hi = {mx(ahi.x, bhi.x), mx(ahi.y, bhi.y), mx(ahi.z, bhi.z)};
  const float ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
  const float a = ex * ey + ey * ez + ez * ex;
  return (a == a) ? a : std::numeric_limits<float>::infinity();
}

static void ploc(Build& b, uint32_t radius) {
  const uint32_t L = b.L;
  if (L < 2) return;
  const uint32_t Lm1 = L - 1;
  struct Cl {
    uint32_t node;
    f3 lo, hi;
  };
  std::vector<Cl> cur(L), nxt;
  for (uint32_t m = 0; m < L; m++) cur[m] = {Lm1 + m, b.lr[Lm1 + m].bmin, b.lr[Lm1 + m].bmax};
  std::vector<uint32_t> nn;
  uint32_t next_id = L - 2;
  while (cur.size() > 1) {
    const uint32_t n = uint32_t(cur.size());
    nn.assign(n, 0);
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t lo = i > radius ? i - radius : 0, hi = std::min(n - 1, i + radius);
      float best = 0.0f;
      uint32_t bj = 0xFFFFFFFFu;
      for (uint32_t j = lo; j <= hi; j++) {
        if (j == i) continue;
        const float a = union_half_area(cur[i].lo, cur[i].hi, cur[j].lo, cur[j].hi);
        if (bj == 0xFFFFFFFFu || a < best) {
          best = a;
          bj = j;
        }
      }
      nn[i] = bj;
    }
    nxt.clear();
    uint32_t merged = 0;
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t j = nn[i];
      const bool mutual = nn[j] == i;
      if (mutual && i > j) continue;  // merged into j's node
      if (!mutual) {
        nxt.push_back(cur[i]);
        continue;
      }
      const uint32_t id = next_id - merged++;
      Node nd;
      nd.bmin = {mn(cur[i].lo.x, cur[j].lo.x), mn(cur[i].lo.y, cur[j].lo.y), mn(cur[i].lo.z, cur[j].lo.z)};
      nd.bmax = {mx(cur[i].hi.x, cur[j].hi.x), mx(cur[i].hi.y, cur[j].hi.y), mx(cur[i].hi.z, cur[j].hi.z)};
      nd.left = cur[i].node;
      nd.escape = cur[j].node;
      b.lr[id] = nd;
      nxt.push_back({id, nd.bmin, nd.bmax});
    }
    next_id -= merged;
    cur.swap(nxt);
  }
  // subtree counts for the peach layout (parents have smaller ids than their children)
  for (uint32_t i = Lm1; i-- > 0;)
    b.temp_count[i] = b.temp_count[b.lr[i].left] + b.temp_count[b.lr[i].escape] + 2;
  b.peach.clear();
  build_peach(b);
  b.q32.clear();
  b.q16.clear();
}

// ---------------------------------------------------------------------------
// BVH4: collapse of the Karras BVH2, FP16 boxes, traversal
// ---------------------------------------------------------------------------
// Collapse rule (core/builders/cbvh_core.cpp:263-352,
// ForceGet4ChildernForDynamicLRFormat): a BVH4 node replaces a BVH2 internal
// node X; its children are X's children with every internal child replaced by
// its own two children, in left-to-right order (the reference sorts the
// gathered children by the first leaf they cover, which is that order).
// Unused slots get box (+1e9, -1e9) and the empty leaf PackOffsetAndSize(0,0)
// (cbvh_core.cpp:402-409 gives the box).  BVH4 nodes are therefore the BVH2
// internal nodes at even depth; they are numbered in Karras index order.

static void collapse4(const Build& b, std::vector<Quad>& quads) {
  quads.clear();
  const uint32_t L = b.L;
  const Node* lr = b.lr.data();
  auto empty_child = []() {
    Child4 e;
    e.bmin = {1e9f, 1e9f, 1e9f};
    e.bmax = {-1e9f, -1e9f, -1e9f};
    e.ref = pack_leaf(0, 0);
    return e;
  };
  if (L == 1) {  // root is a leaf: one node holding it and the peach dummy
    Quad q;
    q.c[0] = Child4{lr[0].bmin, lr[0].bmax, lr[0].left};
    q.c[1] = Child4{{0, 0, 0}, {0, 0, 0}, 0x80000000u};
    q.c[2] = empty_child();
    q.c[3] = empty_child();
    quads.push_back(q);
    return;
  }
  // depth of every internal node
  std::vector<int> depth(L - 1, -1);
  depth[0] = 0;
  std::vector<uint32_t> st = {0};
  while (!st.empty()) {
    uint32_t k = st.back();
    st.pop_back();
    for (uint32_t c : {lr[k].left, lr[k].escape})
      if (c < L - 1) { depth[c] = depth[k] + 1; st.push_back(c); }
  }
  std::vector<uint32_t> qidx(L - 1, 0xFFFFFFFFu);
  uint32_t nq = 0;
  for (uint32_t k = 0; k + 1 < L; k++)
    if (depth[k] % 2 == 0) qidx[k] = nq++;
  quads.resize(nq);
  // LR index -> child record (leaf refs are the packed leaves, internal refs the quad index)
  auto child_of = [&](uint32_t idx) {
    Child4 c;
    c.bmin = lr[idx].bmin;
    c.bmax = lr[idx].bmax;
    c.ref = (idx >= L - 1) ? lr[idx].left : qidx[idx];
    return c;
  };
  for (uint32_t k = 0; k + 1 < L; k++) {
    if (qidx[k] == 0xFFFFFFFFu) continue;
    Quad& q = quads[qidx[k]];
    int n = 0;
    for (uint32_t c : {lr[k].left, lr[k].escape}) {
      if (c < L - 1) {
        q.c[n++] = child_of(lr[c].left);
        q.c[n++] = child_of(lr[c].escape);
      } else {
        q.c[n++] = child_of(c);
      }
    }
    while (n < 4) q.c[n++] = empty_child();
  }
}

// float -> half, round to nearest even (IEEE 754 binary16), as half.hpp's
// default conversion (round_to_nearest, half.hpp:820-834) does.
static uint16_t f2h_rne(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7FFFFFFFu;
  if (ax > 0x7F800000u) return uint16_t(sign | 0x7E00u | ((ax >> 13) & 0x3FFu));  // NaN
  if (ax >= 0x477FF000u) return uint16_t(sign | 0x7C00u);  // >= 65520 rounds to infinity
  const uint32_t e = ax >> 23;
  if (e >= 113) {  // normal half
    const uint32_t r = ax - 0x38000000u;
    uint32_t h = r >> 13;
    const uint32_t rem = r & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return uint16_t(sign | h);
  }
  // subnormal half (or zero): value / 2^-24 rounded to nearest even
  const uint32_t m = (ax & 0x7FFFFFu) | (e ? 0x800000u : 0u);
  const uint32_t shift = (e ? 126u - e : 125u);
  if (shift > 31) return uint16_t(sign);
  uint32_t q = m >> shift;
  const uint32_t rem = m & ((1u << shift) - 1u);
  const uint32_t half = 1u << (shift - 1u);
  if (rem > half || (rem == half && (q & 1u))) q++;
  return uint16_t(sign | q);
}

// nextafter(from, to) on binary16 (the step half.hpp's nextafter takes), for
// the two targets the packing uses (+-65504): the half's value as a signed
// integer key (+0 and -0 both 0, -x = -key(x); keys are in value order),
// one step toward the target's key, back to a half.  A step onto zero keeps
// the sign of `from`; a step off zero lands on the smallest subnormal of the
// target's sign; the target itself and NaN stay.
static uint16_t h_nextafter(uint16_t from, uint16_t to) {
  if ((from & 0x7FFFu) > 0x7C00u) return from;
  if (from == to) return to;
  auto key = [](uint16_t h) { return (h & 0x8000u) ? -int32_t(h & 0x7FFFu) : int32_t(h & 0x7FFFu); };
  const int32_t kf = key(from), kt = key(to);
  if (kf == kt) return to;  // the other zero
  const int32_t k = kf + (kt > kf ? 1 : -1);
  if (k == 0) return uint16_t(from & 0x8000u);
  return k > 0 ? uint16_t(k) : uint16_t(0x8000u | uint32_t(-k));
}
static uint16_t round_to_neg(float v) { return h_nextafter(f2h_rne(v), 0xFBFFu); }  // cbvh_fat.cpp:447
static uint16_t round_to_pos(float v) { return h_nextafter(f2h_rne(v), 0x7BFFu); }  // cbvh_fat.cpp:448

static float h2f(uint16_t h) {
  const uint32_t sign = uint32_t(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
  uint32_t x;
  if (e == 0) {
    if (m == 0) x = sign;
    else {  // subnormal: m * 2^-24
      float f = float(m) * (1.0f / 16777216.0f);
      std::memcpy(&x, &f, 4);
      x |= sign;
    }
  } else if (e == 31) x = sign | 0x7F800000u | (m << 13);
  else x = sign | ((e + 112u) << 23) | (m << 13);
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

// FP16 child: PackNode (cbvh_fat.cpp:477-497) then UnpackBVHNode
// (BVH4HalfRT.cpp:56-71) -- the FP32 box the traversal actually tests.
static Child4 fp16_child(const Child4& c) {
  Child4 r;
  r.bmin = {h2f(round_to_neg(c.bmin.x)), h2f(round_to_neg(c.bmin.y)), h2f(round_to_neg(c.bmin.z))};
  r.bmax = {h2f(round_to_pos(c.bmax.x)), h2f(round_to_pos(c.bmax.y)), h2f(round_to_pos(c.bmax.z))};
  r.ref = c.ref;
  return r;
}

// Nearest-hit rule of the BVH4 path: (t, original primId) lexicographic, so
// the answer does not depend on the visiting order (FP16 and FP32 nodes
// visit in different orders).  Any-hit stops at the first accepted triangle.
static inline bool mt_test_lex(f3 o, f3 d, float tNear, f3 A, f3 B, f3 C, Hit& h, uint32_t prim) {
  float t, u, v;
  // t == tFar with no hit yet is outside the open interval (tNear, tFar)
  if (mt_eval(o, d, tNear, A, B, C, t, u, v) &&
      (t < h.t || (t == h.t && h.primId != 0xFFFFFFFFu && prim < h.primId))) {
    h.t = t;
    h.primId = prim;
    h.instId = 0;
    h.geomId = 0;
    h.u = u;
    h.v = v;
    return true;
  }
  return false;
}

// BVH4HalfRT::BVH4Traverse order (BVH4HalfRT.cpp:73-256): four slab tests,
// misses get MAXFLOAT, five-comparator network on the entry distances, push
// w, z, y (if hit), continue with x.
// fused: the box tests in slab_fused form (the engine's any-hit calls on FP16 nodes).
static Hit trace_bvh4(const Build& b, const std::vector<Quad>& quads, f4 posNear, f4 dirFar, Counters* cnt,
                      bool fused = false) {
  const bool anyhit = (dirFar.w <= 0.0f);
  const float tFar = anyhit ? -dirFar.w : dirFar.w;
  Hit h;
  h.t = tFar;
  h.primId = h.instId = h.geomId = 0xFFFFFFFFu;
  h.u = h.v = 0.0f;
  const f3 o = {posNear.x, posNear.y, posNear.z};
  const f3 d = {dirFar.x, dirFar.y, dirFar.z};
  const float tNear = posNear.w;
  const f3 inv = safe_inverse(d);
  // the watertight mode's robust box test replaces the fused one
  const bool use_fused = fused && fused_ok(o, inv) && g_tri_test != 1;
  uint32_t stack[64];
  int top = 0;
  uint32_t cur = 0;
  const float MAXF = 3.402823466e+38f;
  while (top >= 0) {
    bool pop = false;
    if ((cur & LEAF_BIT) == 0) {
      const Quad& q = quads[cur];
      if (cnt) cnt->pairs++;
      float tm[4];
      uint32_t ch[4];
      for (int k = 0; k < 4; k++) {
        float a, c;
        if (use_fused) slab_fused(o, inv, q.c[k].bmin, q.c[k].bmax, a, c);
        else slab(o, d, inv, q.c[k].bmin, q.c[k].bmax, a, c);
        // An unused slot (+1e9, -1e9) is inside out; the min/max slab test
        // turns it into the box [-1e9, 1e9], so the reference enters it and
        // finds its empty leaf.  Skipping slots whose reference is the empty
        // leaf PackOffsetAndSize(0, 0) gives the same hits without the visit.
        const bool hit = (a <= c) && (c >= tNear) && (a <= h.t) && q.c[k].ref != 0x80000000u;
        tm[k] = hit ? a : MAXF;
        ch[k] = q.c[k].ref;
      }
      auto cswap = [&](int i, int j) {  // swap when tm[j] < tm[i]
        if (tm[j] < tm[i]) { std::swap(tm[i], tm[j]); std::swap(ch[i], ch[j]); }
      };
      cswap(0, 1);
      cswap(2, 3);
      cswap(0, 2);
      cswap(1, 3);
      cswap(1, 2);
      if (tm[3] < MAXF) stack[top++] = ch[3];
      if (tm[2] < MAXF) stack[top++] = ch[2];
      if (tm[1] < MAXF) stack[top++] = ch[1];
      if (tm[0] < MAXF) cur = ch[0];
      pop = !(tm[0] < MAXF);
      if (cnt && uint64_t(top) > cnt->max_stack) cnt->max_stack = uint64_t(top);
    }
    if (!pop && (cur & LEAF_BIT) != 0 && cur != 0xFFFFFFFFu) {
      const uint32_t start = leaf_start(cur), count = leaf_count(cur);
      if (cnt) { cnt->leaves++; cnt->tris += count; }
      for (uint32_t t = start; t < start + count; t++) {
        const uint32_t* ix = &b.idx_leaf_order[3 * size_t(t)];
        if (mt_test_lex(o, d, tNear, vpos(b, ix[0]), vpos(b, ix[1]), vpos(b, ix[2]), h, b.perm[t]) && anyhit) break;
      }
      pop = true;
      if (anyhit && h.primId != 0xFFFFFFFFu) break;
    }
    if (pop) {
      top--;
      cur = stack[std::max(top, 0)];
    }
  }
  return h;
}

// Brute force with the (t, primId) rule.
static Hit trace_brute_lex(const Build& b, f4 posNear, f4 dirFar) {
  const bool anyhit = (dirFar.w <= 0.0f);
  Hit h;
  h.t = anyhit ? -dirFar.w : dirFar.w;
  h.primId = h.instId = h.geomId = 0xFFFFFFFFu;
  h.u = h.v = 0.0f;
  const f3 o = {posNear.x, posNear.y, posNear.z};
  const f3 d = {dirFar.x, dirFar.y, dirFar.z};
  for (uint32_t t = 0; t < b.ntris; t++)
    mt_test_lex(o, d, posNear.w, vpos(b, b.idx[3 * t]), vpos(b, b.idx[3 * t + 1]), vpos(b, b.idx[3 * t + 2]), h, t);
  return h;
}

// ---------------------------------------------------------------------------
// Instances and the two-level traversal
// (BVH2CommonRT::AddInstance / CommitScene / RayQuery_NearestHit,
//  BVH2CommonRT_host.cpp:118-164, BVH2CommonRT.cpp:143-279)
// ---------------------------------------------------------------------------
struct Inst {
  uint32_t geom;
  float fwd[16];   // column-major, as AddInstance receives it
  float inv[12];   // rows 0..2 of the inverse (world -> object)
  f4 bmin, bmax;   // world box
};

struct Scene {
  std::vector<Build*> geoms;
  std::vector<Inst> inst;
  Build tlas;
  ~Scene() {
    for (Build* g : geoms) delete g;
  }
};

// LiteMath float4x4 * float4 for a column-major matrix, summed left to right:
// r.k = ((m[0+k]*x + m[4+k]*y) + m[8+k]*z) + m[12+k]*w.
static inline f4 mat_vec(const float* m, f4 v) {
  return {((m[0] * v.x + m[4] * v.y) + m[8] * v.z) + m[12] * v.w,
          ((m[1] * v.x + m[5] * v.y) + m[9] * v.z) + m[13] * v.w,
          ((m[2] * v.x + m[6] * v.y) + m[10] * v.z) + m[14] * v.w,
          ((m[3] * v.x + m[7] * v.y) + m[11] * v.z) + m[15] * v.w};
}

// 4x4 inverse by the Laplace expansion in 2x2 minors (D. Eberly, "The Laplace
// Expansion Theorem: Computing the Determinants and Inverses of Matrices",
// Geometric Tools), evaluated in double and rounded to float once per entry.
// Writes rows 0..2; returns false for a singular or non-finite matrix.
static bool inverse_rows(const float* m, float* out12) {
  double a[4][4];
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) a[r][c] = double(m[4 * c + r]);
  const double s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
  const double s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
  const double s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
  const double s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
  const double s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
  const double s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
  const double c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
  const double c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
  const double c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
  const double c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
  const double c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
  const double c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
  const double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  if (det == 0.0 || !std::isfinite(det)) return false;
  const double b[12] = {
      a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3,   -a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3,
      a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3,   -a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3,
      -a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1,  a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1,
      -a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1,  a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1,
      a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0,   -a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0,
      a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0,   -a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0};
  for (int k = 0; k < 12; k++) {
    out12[k] = float(b[k] / det);
    if (!std::isfinite(out12[k])) return false;
  }
  return true;
}

// matmul4x3 / matmul3x3 (raytrace_common.h:82-90) with the inverse rows.
static inline f3 xform_point(const float* r, f3 p) {
  return {((r[0] * p.x + r[1] * p.y) + r[2] * p.z) + r[3] * 1.0f,
          ((r[4] * p.x + r[5] * p.y) + r[6] * p.z) + r[7] * 1.0f,
          ((r[8] * p.x + r[9] * p.y) + r[10] * p.z) + r[11] * 1.0f};
}
static inline f3 xform_dir(const float* r, f3 d) {
  return {((r[0] * d.x + r[1] * d.y) + r[2] * d.z) + r[3] * 0.0f,
          ((r[4] * d.x + r[5] * d.y) + r[6] * d.z) + r[7] * 0.0f,
          ((r[8] * d.x + r[9] * d.y) + r[10] * d.z) + r[11] * 0.0f};
}

// Instance box: the box of the eight transformed corners of the geometry's
// vertex box (BVH2CommonRT_host.cpp:133-157).
static void instance_box(const Build& g, const float* fwd, f4& lo, f4& hi) {
  const f4 a = g.bbmin, b = g.bbmax;
  const f4 corners[8] = {{a.x, a.y, a.z, 1.0f}, {b.x, a.y, a.z, 1.0f}, {a.x, b.y, a.z, 1.0f},
                         {a.x, a.y, b.z, 1.0f}, {b.x, b.y, a.z, 1.0f}, {b.x, a.y, b.z, 1.0f},
                         {a.x, b.y, b.z, 1.0f}, {b.x, b.y, b.z, 1.0f}};
  for (int k = 0; k < 8; k++) {
    const f4 p = mat_vec(fwd, corners[k]);
    if (k == 0) {
      lo = p;
      hi = p;
    } else {
      lo = {mn(lo.x, p.x), mn(lo.y, p.y), mn(lo.z, p.z), mn(lo.w, p.w)};
      hi = {mx(hi.x, p.x), mx(hi.y, p.y), mx(hi.z, p.z), mx(hi.w, p.w)};
    }
  }
}

// Two-level acceptance: inside one instance the first strictly closer hit
// wins; at equal distance a smaller instance id replaces a larger one.
static inline bool mt_test_inst(f3 o, f3 d, float tNear, f3 A, f3 B, f3 C, Hit& h, uint32_t prim, uint32_t inst,
                                uint32_t geom) {
  float t, u, v;
  if (mt_eval(o, d, tNear, A, B, C, t, u, v) &&
      (t < h.t || (t == h.t && h.primId != 0xFFFFFFFFu && inst < h.instId))) {
    h.t = t;
    h.primId = prim;
    h.instId = inst;
    h.geomId = geom;
    h.u = u;
    h.v = v;
    return true;
  }
  return false;
}

static inline uint32_t lr_ref(const Build& t, uint32_t node) {  // child reference as the GPU pairs store it
  return (node >= t.L - 1) ? t.lr[node].left : node;
}

// Nearest / any hit over all instances.  The top-level tree is the LBVH over
// instance boxes, walked with the same near-first rule as the BLAS (left child
// first on equal entry distance, culling with the current hit distance); at a
// top-level leaf every instance of the leaf is entered in leaf order with the
// ray moved into object space, and the BLAS is walked with the running hit.
static Hit trace_scene(const Scene& s, f4 posNear, f4 dirFar, Counters* cnt) {
  const bool anyhit = (dirFar.w <= 0.0f);
  Hit h;
  h.t = anyhit ? -dirFar.w : dirFar.w;
  h.primId = h.instId = h.geomId = 0xFFFFFFFFu;
  h.u = h.v = 0.0f;
  if (s.inst.empty()) return h;
  const f3 wo = {posNear.x, posNear.y, posNear.z};
  const f3 wd = {dirFar.x, dirFar.y, dirFar.z};
  const float tNear = posNear.w;
  const f3 winv = safe_inverse(wd);
  const Build& T = s.tlas;
  auto visit_leaf = [&](uint32_t ref) -> bool {
    const uint32_t start = leaf_start(ref), count = leaf_count(ref);
    for (uint32_t k = start; k < start + count; k++) {
      const uint32_t id = T.perm[k];
      const Inst& I = s.inst[id];
      const f3 o = xform_point(I.inv, wo);
      const f3 d = xform_dir(I.inv, wd);
      const bool stop = visit_bvh2(*s.geoms[I.geom], o, d, tNear, anyhit, h, cnt,
                                   [&](f3 A, f3 B, f3 C, uint32_t prim) {
                                     mt_test_inst(o, d, tNear, A, B, C, h, prim, id, I.geom);
                                   });
      if (stop) return true;
    }
    return false;
  };
  auto box_hit = [&](const Node& n, float& tmin) {
    float a, b;
    slab(wo, wd, winv, n.bmin, n.bmax, a, b);
    tmin = a;
    return (a <= b) && (b >= tNear) && (a <= h.t);
  };
  if (T.L == 1) {
    // the root is a leaf: one pair of (leaf, empty box)
    if (cnt) cnt->pairs++;
    float a;
    if (box_hit(T.lr[0], a)) visit_leaf(T.lr[0].left);
    return h;
  }
  uint32_t stack[80];
  int top = 0;
  uint32_t cur = 0;
  while (true) {
    bool pop = false;
    if ((cur & LEAF_BIT) == 0) {
      if (cnt) cnt->pairs++;
      const Node& n = T.lr[cur];
      float a0, a1;
      const bool h0 = box_hit(T.lr[n.left], a0);
      const bool h1 = box_hit(T.lr[n.escape], a1);
      const uint32_t c0 = lr_ref(T, n.left), c1 = lr_ref(T, n.escape);
      if (h0 && h1) {
        cur = (a0 <= a1) ? c0 : c1;
        stack[top++] = (a0 <= a1) ? c1 : c0;
      } else if (h0 || h1) {
        cur = h0 ? c0 : c1;
      } else {
        pop = true;
      }
    } else {
      if (visit_leaf(cur)) break;
      pop = true;
    }
    if (pop) {
      if (top == 0) break;
      cur = stack[--top];
    }
  }
  return h;
}

// The closed form the two-level walk must agree with: every instance traced on
// its own (the single-instance traversal with the ray in object space), then
// the smallest t, the smallest instance id on equal t.
static Hit trace_scene_per_instance(const Scene& s, f4 posNear, f4 dirFar) {
  const bool anyhit = (dirFar.w <= 0.0f);
  Hit best;
  best.t = anyhit ? -dirFar.w : dirFar.w;
  best.primId = best.instId = best.geomId = 0xFFFFFFFFu;
  best.u = best.v = 0.0f;
  for (uint32_t id = 0; id < s.inst.size(); id++) {
    const Inst& I = s.inst[id];
    const f3 o = xform_point(I.inv, f3{posNear.x, posNear.y, posNear.z});
    const f3 d = xform_dir(I.inv, f3{dirFar.x, dirFar.y, dirFar.z});
    Hit r = trace_bvh2common(*s.geoms[I.geom], f4{o.x, o.y, o.z, posNear.w}, f4{d.x, d.y, d.z, dirFar.w}, nullptr);
    if (r.primId != 0xFFFFFFFFu && r.t < best.t) {
      best = r;
      best.instId = id;
      best.geomId = I.geom;
    }
  }
  return best;
}

// Brute force over every triangle of every instance (ascending instance id,
// then ascending triangle id, first strictly closer hit wins).
static Hit trace_scene_brute(const Scene& s, f4 posNear, f4 dirFar) {
  const bool anyhit = (dirFar.w <= 0.0f);
  Hit h;
  h.t = anyhit ? -dirFar.w : dirFar.w;
  h.primId = h.instId = h.geomId = 0xFFFFFFFFu;
  h.u = h.v = 0.0f;
  const float tNear = posNear.w;
  for (uint32_t id = 0; id < s.inst.size(); id++) {
    const Inst& I = s.inst[id];
    const Build& g = *s.geoms[I.geom];
    const f3 o = xform_point(I.inv, f3{posNear.x, posNear.y, posNear.z});
    const f3 d = xform_dir(I.inv, f3{dirFar.x, dirFar.y, dirFar.z});
    for (uint32_t t = 0; t < g.ntris; t++) {
      float tt, u, v;
      if (mt_eval(o, d, tNear, vpos(g, g.idx[3 * t]), vpos(g, g.idx[3 * t + 1]), vpos(g, g.idx[3 * t + 2]), tt, u,
                  v) &&
          tt < h.t) {
        h.t = tt;
        h.primId = t;
        h.instId = id;
        h.geomId = I.geom;
        h.u = u;
        h.v = v;
      }
    }
  }
  return h;
}

}  // namespace

// ===========================================================================
// C entry points (loaded by tests/ and bench.py through ctypes)
// ===========================================================================
namespace {
std::vector<Quad>& quads_for(const Build* cb, int fp16) {
  Build* b = const_cast<Build*>(cb);
  if (b->q32.empty()) {
    collapse4(*b, b->q32);
    b->q16 = b->q32;
    for (auto& q : b->q16)
      for (auto& c : q.c) c = fp16_child(c);
  }
  return fp16 ? b->q16 : b->q32;
}
}  // namespace

extern "C" {

void* hrto_build_ex(const float* vert4, uint32_t nVert, const uint32_t* idx, uint32_t nIdx, int cubic) {
  if (nIdx < 3) return nullptr;
  Build* b = new Build();
  b->cubic = cubic != 0;
  b->verts.resize(nVert);
  std::memcpy(b->verts.data(), vert4, sizeof(f4) * nVert);
  for (auto& v : b->verts) v.w = 1.0f;  // BVH2CommonRT_host.cpp:81 stores w = 1
  b->ntris = nIdx / 3;
  b->idx.assign(idx, idx + 3 * size_t(b->ntris));
  build_all(*b);
  return b;
}

void* hrto_build(const float* vert4, uint32_t nVert, const uint32_t* idx, uint32_t nIdx) {
  return hrto_build_ex(vert4, nVert, idx, nIdx, 0);
}

void hrto_free(void* h) { delete static_cast<Build*>(h); }

// Replaces the hierarchy above the leaves with the agglomerative one (ploc);
// leaves, leaf order and triangles stay.
void hrto_ploc(void* h, uint32_t radius) { ploc(*static_cast<Build*>(h), radius); }

// Refit (BuildQuality BUILD_REFIT, CrossRT.h:14): new vertex positions (and
// optionally indices, same triangle count), hierarchy and leaf order kept;
// leaf boxes are recomputed from the leaf-order triangles and the internal
// boxes by the same 32-sweep refit as a build (lbvh.cpp:443-489).  The
// vertex box is recomputed; the Morton scale, codes and leaf order stay those
// of the original build.  Returns 0, or -1 if the triangle count differs.
int hrto_refit(void* h, const float* vert4, uint32_t nVert, const uint32_t* idx, uint32_t nIdx) {
  Build& b = *static_cast<Build*>(h);
  if (idx && nIdx / 3 != b.ntris) return -1;
  b.verts.resize(nVert);
  std::memcpy(b.verts.data(), vert4, sizeof(f4) * nVert);
  for (auto& v : b.verts) v.w = 1.0f;
  if (idx) b.idx.assign(idx, idx + 3 * size_t(b.ntris));
  f4 lo = {LBVH_INF, LBVH_INF, LBVH_INF, LBVH_INF}, hi = {-LBVH_INF, -LBVH_INF, -LBVH_INF, -LBVH_INF};
  for (const f4& v : b.verts) {
    lo = {mn(lo.x, v.x), mn(lo.y, v.y), mn(lo.z, v.z), mn(lo.w, v.w)};
    hi = {mx(hi.x, v.x), mx(hi.y, v.y), mx(hi.z, v.z), mx(hi.w, v.w)};
  }
  b.bbmin = lo;
  b.bbmax = hi;
  const uint32_t L = b.L, N = b.ntris;
  for (uint32_t i = 0; i < L; i++) {
    const uint32_t start = b.run_start[i];
    const uint32_t count = (i < L - 1) ? b.run_start[i + 1] - start : N - b.run_start[L - 1];
    f4 bl = {LBVH_INF, LBVH_INF, LBVH_INF, 0}, bh = {-LBVH_INF, -LBVH_INF, -LBVH_INF, 0};
    for (uint32_t j = start; j < start + count; j++) {
      const uint32_t t = b.perm[j];
      const f4 v1 = b.verts[b.idx[3 * t + 0]], v2 = b.verts[b.idx[3 * t + 1]], v3 = b.verts[b.idx[3 * t + 2]];
      bl = {mn(bl.x, mn(v1.x, mn(v2.x, v3.x))), mn(bl.y, mn(v1.y, mn(v2.y, v3.y))), mn(bl.z, mn(v1.z, mn(v2.z, v3.z))), 0};
      bh = {mx(bh.x, mx(v1.x, mx(v2.x, v3.x))), mx(bh.y, mx(v1.y, mx(v2.y, v3.y))), mx(bh.z, mx(v1.z, mx(v2.z, v3.z))), 0};
    }
    Node& nd = b.lr[L - 1 + i];
    nd.bmin = {bl.x, bl.y, bl.z};
    nd.bmax = {bh.x, bh.y, bh.z};
  }
  refit(b);
  b.peach.clear();
  build_peach(b);
  b.idx_leaf_order.resize(3 * size_t(N));
  for (uint32_t i = 0; i < N; i++)
    for (int k = 0; k < 3; k++) b.idx_leaf_order[3 * i + k] = b.idx[3 * b.perm[i] + k];
  b.q32.clear();
  b.q16.clear();
  return 0;
}

// sizes: [ntris, leaves, lr_nodes, peach_nodes, max_run]
void hrto_sizes(void* h, uint64_t* out5) {
  const Build* b = static_cast<Build*>(h);
  out5[0] = b->ntris;
  out5[1] = b->L;
  out5[2] = b->lr.size();
  out5[3] = b->peach.size();
  out5[4] = b->max_run;
}

// scene box (8 floats: min xyzw, max xyzw) and Morton scale (4 floats)
void hrto_scene_box(void* h, float* out12) {
  const Build* b = static_cast<Build*>(h);
  std::memcpy(out12 + 0, &b->bbmin, 16);
  std::memcpy(out12 + 4, &b->bbmax, 16);
  std::memcpy(out12 + 8, &b->zscale, 16);
}

void hrto_codes(void* h, uint32_t* codes_unsorted, uint32_t* sorted_codes, uint32_t* sorted_ids) {
  const Build* b = static_cast<Build*>(h);
  std::memcpy(codes_unsorted, b->codes_unsorted.data(), 4 * size_t(b->ntris));
  std::memcpy(sorted_codes, b->sorted_codes.data(), 4 * size_t(b->ntris));
  std::memcpy(sorted_ids, b->sorted_ids.data(), 4 * size_t(b->ntris));
}

void hrto_runs(void* h, uint32_t* run_start, uint32_t* ucodes) {
  const Build* b = static_cast<Build*>(h);
  std::memcpy(run_start, b->run_start.data(), 4 * size_t(b->L));
  std::memcpy(ucodes, b->ucodes.data(), 4 * size_t(b->L));
}

// Karras left/right layout: 8 words per node (bmin xyz, left, bmax xyz, right)
void hrto_lr_nodes(void* h, uint32_t* out) {
  const Build* b = static_cast<Build*>(h);
  std::memcpy(out, b->lr.data(), sizeof(Node) * b->lr.size());
}

void hrto_peach_nodes(void* h, uint32_t* out) {
  const Build* b = static_cast<Build*>(h);
  std::memcpy(out, b->peach.data(), sizeof(Node) * b->peach.size());
}

void hrto_perm(void* h, uint32_t* out) {
  const Build* b = static_cast<Build*>(h);
  std::memcpy(out, b->perm.data(), 4 * size_t(b->ntris));
}

void hrto_temp_count(void* h, uint32_t* out) {
  const Build* b = static_cast<Build*>(h);
  std::memcpy(out, b->temp_count.data(), 4 * b->temp_count.size());
}

// mode: 0 = BVH2Common over the peach layout, 1 = left/right layout,
//       2 = brute force.  rays: 8 floats each (posNear, dirFar).
// hits: 6 words each (t, primId, instId, geomId, u, v).
// counters (mode 0 only, may be null): per ray [pairs, tris].
void hrto_trace(void* h, int mode, const float* rays, uint64_t n, uint32_t* hits, uint32_t* counters, int threads) {
  const Build* b = static_cast<Build*>(h);
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256) if (threads != 1)
#endif
  for (int64_t i = 0; i < int64_t(n); i++) {
    f4 p, d;
    std::memcpy(&p, rays + 8 * i, 16);
    std::memcpy(&d, rays + 8 * i + 4, 16);
    Hit r;
    Counters c;
    if (mode == 0) r = trace_bvh2common(*b, p, d, counters ? &c : nullptr);
    else if (mode == 1) r = trace_lr(*b, p, d);
    else r = trace_brute(*b, p, d);
    std::memcpy(hits + 6 * i, &r, sizeof(Hit));
    if (counters && mode == 0) {
      counters[2 * i + 0] = uint32_t(c.pairs);
      counters[2 * i + 1] = uint32_t(c.tris);
    }
  }
}

// FP64 Moller-Trumbore t for a given triangle (t-error reporting).
// Deepest traversal stack per ray (mode 0: BVH2 peach order, 1: BVH4 FP32).
void hrto_max_stack(void* h, int mode, const float* rays, uint64_t n, uint32_t* out, int threads) {
  const Build* b = static_cast<Build*>(h);
  const std::vector<Quad>* q = (mode == 1) ? &quads_for(b, 0) : nullptr;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256) if (threads != 1)
#endif
  for (int64_t i = 0; i < int64_t(n); i++) {
    f4 p, d;
    std::memcpy(&p, rays + 8 * i, 16);
    std::memcpy(&d, rays + 8 * i + 4, 16);
    Counters c;
    if (q) (void)trace_bvh4(*b, *q, p, d, &c);
    else (void)trace_bvh2common(*b, p, d, &c);
    out[i] = uint32_t(c.max_stack);
  }
}

void hrto_t_fp64(void* h, const float* rays, const uint32_t* prim, uint64_t n, double* out_t) {
  const Build* b = static_cast<Build*>(h);
  for (uint64_t i = 0; i < n; i++) {
    if (prim[i] == 0xFFFFFFFFu) { out_t[i] = NAN; continue; }
    const uint32_t* ix = &b->idx[3 * size_t(prim[i])];
    double A[3], B[3], C[3], o[3], d[3];
    for (int k = 0; k < 3; k++) {
      A[k] = (&b->verts[ix[0]].x)[k];
      B[k] = (&b->verts[ix[1]].x)[k];
      C[k] = (&b->verts[ix[2]].x)[k];
      o[k] = rays[8 * i + k];
      d[k] = rays[8 * i + 4 + k];
    }
    double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    double e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    double tv[3] = {o[0] - A[0], o[1] - A[1], o[2] - A[2]};
    double pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    double qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
    double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    out_t[i] = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) / det;
  }
}

// BVH4 nodes: quads[i] = 4 children x 8 words (bmin xyz, ref, bmax xyz, 0);
// fp16 = 1 gives the boxes after FP16 packing and unpacking.
uint64_t hrto_quad_count(void* h) { return quads_for(static_cast<Build*>(h), 0).size(); }
void hrto_quads(void* h, int fp16, uint32_t* out) {
  const std::vector<Quad>& q = quads_for(static_cast<Build*>(h), fp16);
  for (size_t i = 0; i < q.size(); i++)
    for (int k = 0; k < 4; k++) {
      uint32_t* w = out + 32 * i + 8 * k;
      std::memcpy(w + 0, &q[i].c[k].bmin, 12);
      w[3] = q[i].c[k].ref;
      std::memcpy(w + 4, &q[i].c[k].bmax, 12);
      w[7] = 0;
    }
}

// BVH4 traversal (mode 0: FP32 boxes, 1: FP16 boxes, 3: FP16 boxes with
// slab_fused box tests) and brute force with the (t, primId) rule (mode 2).
void hrto_trace4(void* h, int mode, const float* rays, uint64_t n, uint32_t* hits, uint32_t* counters, int threads) {
  const Build* b = static_cast<Build*>(h);
  const std::vector<Quad>* q = (mode == 2) ? nullptr : &quads_for(b, mode == 3 ? 1 : mode);
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256) if (threads != 1)
#endif
  for (int64_t i = 0; i < int64_t(n); i++) {
    f4 p, d;
    std::memcpy(&p, rays + 8 * i, 16);
    std::memcpy(&d, rays + 8 * i + 4, 16);
    Counters c;
    Hit r = q ? trace_bvh4(*b, *q, p, d, counters ? &c : nullptr, mode == 3) : trace_brute_lex(*b, p, d);
    std::memcpy(hits + 6 * i, &r, sizeof(Hit));
    if (counters && q) {
      counters[2 * i + 0] = uint32_t(c.pairs);
      counters[2 * i + 1] = uint32_t(c.tris);
    }
  }
}

void hrto_set_triangle_test(int test) { g_tri_test = test; }

// BVH2CommonRT traversal on one thread with the reference's metrics
// bookkeeping on (the "as shipped" CPU baseline).  hits as hrto_trace;
// out[0..5]: NC, LC, TC, BLB, SOC, touched 128-byte lines (WSS[0]).
void hrto_trace_metrics(void* h, const float* rays, uint64_t n, uint32_t* hits, uint64_t* out6) {
  const Build* b = static_cast<Build*>(h);
  RefMetrics m;
  for (uint64_t i = 0; i < n; i++) {
    f4 p, d;
    std::memcpy(&p, rays + 8 * i, 16);
    std::memcpy(&d, rays + 8 * i + 4, 16);
    const Hit r = trace_bvh2common(*b, p, d, nullptr, &m);
    std::memcpy(hits + 6 * i, &r, sizeof(Hit));
  }
  out6[0] = m.NC;
  out6[1] = m.LC;
  out6[2] = m.TC;
  out6[3] = m.BLB;
  out6[4] = m.SOC;
  out6[5] = m.WSS[0].size();
}

// One triangle test: out = {t, u, v, accepted (0/1)}; test 0 = Moller-Trumbore, 1 = watertight.
void hrto_tri_eval(int test, const float* o3, const float* d3, float tnear, const float* a3, const float* b3,
                   const float* c3, float* out4) {
  const f3 o{o3[0], o3[1], o3[2]}, d{d3[0], d3[1], d3[2]};
  const f3 A{a3[0], a3[1], a3[2]}, B{b3[0], b3[1], b3[2]}, C{c3[0], c3[1], c3[2]};
  float t = 0.f, u = 0.f, v = 0.f;
  const bool ok = test == 1 ? wt_eval(o, d, tnear, A, B, C, t, u, v) : mt_eval(o, d, tnear, A, B, C, t, u, v);
  out4[0] = t;
  out4[1] = u;
  out4[2] = v;
  out4[3] = ok ? 1.0f : 0.0f;
}

// Brute force: for every ray, how many triangles the test accepts with t in
// (tNear, |tFar|) -- the closed-surface checks count these.
void hrto_tri_hit_counts(int test, const float* v4, uint32_t nverts, const uint32_t* idx, uint32_t ntris,
                         const float* rays, uint64_t n, uint32_t* counts, int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64) if (threads != 1)
#endif
  for (int64_t i = 0; i < int64_t(n); i++) {
    const float* r = rays + 8 * i;
    const f3 o{r[0], r[1], r[2]}, d{r[4], r[5], r[6]};
    const float tfar = std::fabs(r[7]);
    uint32_t c = 0;
    for (uint32_t k = 0; k < ntris; k++) {
      f3 P[3];
      for (int j = 0; j < 3; j++) {
        const uint32_t vi = idx[3 * size_t(k) + j];
        P[j] = vi < nverts ? f3{v4[4 * vi], v4[4 * vi + 1], v4[4 * vi + 2]} : f3{0, 0, 0};
      }
      float t, u, v;
      const bool ok = test == 1 ? wt_eval(o, d, r[3], P[0], P[1], P[2], t, u, v)
                                : mt_eval(o, d, r[3], P[0], P[1], P[2], t, u, v);
      if (ok && t < tfar) c++;
    }
    counts[i] = c;
  }
}

uint32_t hrto_f2h(float v) { return f2h_rne(v); }
uint32_t hrto_round_neg(float v) { return round_to_neg(v); }
uint32_t hrto_round_pos(float v) { return round_to_pos(v); }
float hrto_h2f(uint32_t h) { return h2f(uint16_t(h)); }

// Reference building blocks exposed for known-answer tests.
uint32_t hrto_zindex3d(int x, int y, int z) { return zindex3d(x, y, z); }
uint32_t hrto_pack_leaf(uint32_t start, uint32_t size) { return pack_leaf(start, size); }
void hrto_safe_inverse(const float* d3, float* out3) {
  f3 r = safe_inverse(f3{d3[0], d3[1], d3[2]});
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
void hrto_slab(const float* o3, const float* inv3, const float* bmin3, const float* bmax3, float* out2) {
  slab_ref(f3{o3[0], o3[1], o3[2]}, f3{inv3[0], inv3[1], inv3[2]}, f3{bmin3[0], bmin3[1], bmin3[2]},
       f3{bmax3[0], bmax3[1], bmax3[2]}, out2[0], out2[1]);
}
void hrto_slab_robust(const float* o3, const float* d3, const float* bmin3, const float* bmax3, float* out2) {
  const f3 d{d3[0], d3[1], d3[2]};
  slab_robust(f3{o3[0], o3[1], o3[2]}, d, safe_inverse(d), f3{bmin3[0], bmin3[1], bmin3[2]},
              f3{bmax3[0], bmax3[1], bmax3[2]}, out2[0], out2[1]);
}
void hrto_slab_fused(const float* o3, const float* inv3, const float* bmin3, const float* bmax3, float* out2) {
  slab_fused(f3{o3[0], o3[1], o3[2]}, f3{inv3[0], inv3[1], inv3[2]}, f3{bmin3[0], bmin3[1], bmin3[2]},
       f3{bmax3[0], bmax3[1], bmax3[2]}, out2[0], out2[1]);
}

// ---- scenes with instances --------------------------------------------------
void* hrto_scene_new(void) { return new Scene(); }
void hrto_scene_free(void* h) { delete static_cast<Scene*>(h); }

// Returns the geometry id, or -1 when the index buffer holds no triangle.
int hrto_scene_add_geom(void* h, const float* vert4, uint32_t nVert, const uint32_t* idx, uint32_t nIdx) {
  Scene* s = static_cast<Scene*>(h);
  Build* b = static_cast<Build*>(hrto_build(vert4, nVert, idx, nIdx));
  if (!b) return -1;
  s->geoms.push_back(b);
  return int(s->geoms.size() - 1);
}

// Returns the instance id, or -1 for an unknown geometry or a singular matrix.
int hrto_scene_add_instance(void* h, uint32_t geom, const float* m16) {
  Scene* s = static_cast<Scene*>(h);
  if (geom >= s->geoms.size()) return -1;
  Inst in;
  in.geom = geom;
  std::memcpy(in.fwd, m16, sizeof(in.fwd));
  if (!inverse_rows(m16, in.inv)) return -1;
  instance_box(*s->geoms[geom], in.fwd, in.bmin, in.bmax);
  s->inst.push_back(in);
  return int(s->inst.size() - 1);
}

// Builds the top-level tree; returns its handle (valid for the hrto_* build
// accessors until the scene is freed) or null when there are no instances.
void* hrto_scene_commit(void* h) {
  Scene* s = static_cast<Scene*>(h);
  s->tlas = Build();
  if (s->inst.empty()) return nullptr;
  s->tlas.ntris = uint32_t(s->inst.size());
  for (const Inst& in : s->inst) {
    s->tlas.boxes.push_back(in.bmin);
    s->tlas.boxes.push_back(in.bmax);
  }
  build_all(s->tlas);
  return &s->tlas;
}

// Per instance: 12 floats of inverse rows and 8 floats of world box.
void hrto_scene_inst_info(void* h, float* inv12, float* box8) {
  const Scene* s = static_cast<Scene*>(h);
  for (size_t i = 0; i < s->inst.size(); i++) {
    std::memcpy(inv12 + 12 * i, s->inst[i].inv, 48);
    std::memcpy(box8 + 8 * i, &s->inst[i].bmin, 16);
    std::memcpy(box8 + 8 * i + 4, &s->inst[i].bmax, 16);
  }
}

// mode 0: two-level traversal; 1: per-instance closed form; 2: brute force.
// hits: 6 words per ray; counters (mode 0, may be null): [pairs, tris].
void hrto_scene_trace(void* h, int mode, const float* rays, uint64_t n, uint32_t* hits, uint32_t* counters,
                      int threads) {
  const Scene* s = static_cast<Scene*>(h);
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256) if (threads != 1)
#endif
  for (int64_t i = 0; i < int64_t(n); i++) {
    f4 p, d;
    std::memcpy(&p, rays + 8 * i, 16);
    std::memcpy(&d, rays + 8 * i + 4, 16);
    Counters c;
    Hit r = (mode == 0) ? trace_scene(*s, p, d, counters ? &c : nullptr)
          : (mode == 1) ? trace_scene_per_instance(*s, p, d)
                        : trace_scene_brute(*s, p, d);
    std::memcpy(hits + 6 * i, &r, sizeof(Hit));
    if (counters && mode == 0) {
      counters[2 * i + 0] = uint32_t(c.pairs);
      counters[2 * i + 1] = uint32_t(c.tris);
    }
  }
}


// SuperBlockIndex2DOpt (eye_ray.cpp:114-131): pixel -> thread index.
uint32_t hrto_superblock_offset(uint32_t x, uint32_t y, uint32_t width) {
  const uint32_t inBlockIdX = x & 3u, inBlockIdY = y & 3u;
  const uint32_t localIndex = inBlockIdY * 4 + inBlockIdX;
  const uint32_t wBlocks = width >> 2;
  const uint32_t blockX = x >> 2, blockY = y >> 2;
  const uint32_t inHBlockIdX = blockX & 1u, inHBlockIdY = blockY & 1u;
  const uint32_t localIndexH = inHBlockIdY * 2 + inHBlockIdX;
  const uint32_t wBlocksH = wBlocks >> 1;
  const uint32_t blockHX = blockX >> 1, blockHY = blockY >> 1;
  return (blockHX + blockHY * wBlocksH) * 64 + localIndexH * 16 + localIndex;
}

// kernel_PackXY over every pixel, then kernel_InitEyeRay for every thread
// index.  Matrices are column-major.  rays: 8 floats per thread index.
static uint32_t pcg32_step(uint64_t& state, uint64_t inc) {
  const uint64_t old = state;
  state = old * 6364136223846793005ull + inc;
  const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = uint32_t(old >> 59u);
  return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

// AO rays from primary hits, the engine's ray generator (device code ao.h)
// restated with the same float operations: per pixel the hit point
// o + 0.9999995 * t * d (rtao.cpp:92), the viewer-facing unit geometric
// normal, the OffsRayPos offset (render_common.h:75-112; kGEps = 2e-5);
// per sample k = pixel * spp + s two PCG32 draws (sequence k, seed), the
// cosine-weighted direction (render_common.h:62-70) in Duff et al.'s basis
// about the normal, tNear = 0, tFar = tMax.  A missed pixel gets the empty
// ray (0,0,0,0), (0,1,0,0).  verts3: 3 floats per vertex; hits: compact
// (t bits, primId) per pixel.
// n draws of PCG32 seeded with (initstate = seed, initseq = seq), the
// pcg32_srandom_r / pcg32_random_r pair of O'Neill's pcg-c-basic (the
// generator the AO rays draw from).
void hrto_pcg32(uint64_t seed, uint64_t seq, uint32_t n, uint32_t* out) {
  const uint64_t inc = (seq << 1u) | 1u;
  uint64_t st = 0;
  (void)pcg32_step(st, inc);
  st += seed;
  (void)pcg32_step(st, inc);
  for (uint32_t i = 0; i < n; i++) out[i] = pcg32_step(st, inc);
}

void hrto_ao_rays(const float* verts3, const uint32_t* idx, uint32_t ntris, const float* prim_rays,
                  const uint32_t* hits, uint64_t npix, uint32_t spp, float tmax, uint64_t seed, float* out) {
  const float kGEps = 2e-5f;
  auto step = pcg32_step;
  for (uint64_t pix = 0; pix < npix; pix++) {
    const uint32_t prim = hits[2 * pix + 1];
    float* r = out + 8 * pix * spp;
    if (prim >= ntris) {
      for (uint32_t sIdx = 0; sIdx < spp; sIdx++) {
        const float e[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f};
        std::memcpy(r + 8 * sIdx, e, sizeof(e));
      }
      continue;
    }
    const float* p = prim_rays + 8 * pix;
    float t;
    std::memcpy(&t, &hits[2 * pix], 4);
    const f3 o = {p[0], p[1], p[2]}, d = {p[4], p[5], p[6]};
    const float tt = 0.9999995f * t;
    const f3 hp = {o.x + tt * d.x, o.y + tt * d.y, o.z + tt * d.z};
    const float* A = verts3 + 3 * size_t(idx[3 * size_t(prim) + 0]);
    const float* B = verts3 + 3 * size_t(idx[3 * size_t(prim) + 1]);
    const float* C = verts3 + 3 * size_t(idx[3 * size_t(prim) + 2]);
    const f3 e1 = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    const f3 e2 = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    f3 n = {e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
    const float inv_len = 1.0f / std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
    n = {n.x * inv_len, n.y * inv_len, n.z * inv_len};
    if (n.x * d.x + n.y * d.y + n.z * d.z > 0.0f) n = {-n.x, -n.y, -n.z};
    const float m = std::fmax(std::fmax(std::fabs(hp.x), std::fmax(std::fabs(hp.y), std::fabs(hp.z))), 2.0f * kGEps) * kGEps;
    const f3 org = {hp.x + m * n.x, hp.y + m * n.y, hp.z + m * n.z};
    for (uint32_t sIdx = 0; sIdx < spp; sIdx++) {
      const uint64_t k = pix * spp + sIdx;
      const uint64_t inc = (k << 1u) | 1u;
      uint64_t st = 0;
      (void)step(st, inc);
      st += seed;
      (void)step(st, inc);
      const float r1 = float(step(st, inc) >> 8) * (1.0f / 16777216.0f);
      const float r2 = float(step(st, inc) >> 8) * (1.0f / 16777216.0f);
      const float phi = 6.28318530717958647692f * r1;
      const float sphi = std::sin(phi), cphi = std::cos(phi);
      const float cth = std::sqrt(1.0f - r2);
      const float sth = std::sqrt(r2);
      const float sgn = std::copysign(1.0f, n.z);
      const float a = -1.0f / (sgn + n.z);
      const float b = n.x * n.y * a;
      const f3 tx = {1.0f + sgn * n.x * n.x * a, sgn * b, -sgn * n.x};
      const f3 ty = {b, sgn + n.y * n.y * a, -n.y};
      const float lx = sth * cphi, ly = sth * sphi;
      const float ray[8] = {org.x, org.y, org.z, 0.0f,
                            tx.x * lx + ty.x * ly + n.x * cth, tx.y * lx + ty.y * ly + n.y * cth,
                            tx.z * lx + ty.z * ly + n.z * cth, tmax};
      std::memcpy(r + 8 * sIdx, ray, sizeof(ray));
    }
  }
}

void hrto_eye_rays(const float* projInv, const float* worldViewInv, uint32_t width, uint32_t height, float* rays,
                   uint32_t* packedXY) {
  const size_t n = size_t(width) * height;
  std::vector<uint32_t> xy(n, 0u);
  for (uint32_t y = 0; y < height; y++)
    for (uint32_t x = 0; x < width; x++) xy[hrto_superblock_offset(x, y, width)] = ((y << 16) & 0xFFFF0000u) | (x & 0xFFFFu);
  auto mat_vec = [](const float* a, f4 v) -> f4 {
    return {a[0] * v.x + a[4] * v.y + a[8] * v.z + a[12] * v.w, a[1] * v.x + a[5] * v.y + a[9] * v.z + a[13] * v.w,
            a[2] * v.x + a[6] * v.y + a[10] * v.z + a[14] * v.w, a[3] * v.x + a[7] * v.y + a[11] * v.z + a[15] * v.w};
  };
  auto normalize = [](f3 u) -> f3 {
    const float len = std::sqrt(dot3(u, u));
    return {u.x / len, u.y / len, u.z / len};
  };
  for (size_t t = 0; t < n; t++) {
    const uint32_t XY = xy[t];
    const uint32_t x = XY & 0x0000FFFFu, y = (XY & 0xFFFF0000u) >> 16;
    const float fx = (float(x) + 0.5f) / float(width), fy = (float(y) + 0.5f) / float(height);
    // EyeRayDirNormalized
    f4 pos = {2.0f * fx - 1.0f, -2.0f * fy + 1.0f, 0.0f, 1.0f};
    pos = mat_vec(projInv, pos);
    const float w = pos.w;
    pos = {pos.x / w, pos.y / w, pos.z / w, pos.w / w};
    f3 dir = normalize({pos.x, pos.y, pos.z});
    // transform_ray3f
    const f3 o = {0.0f, 0.0f, 0.0f};
    const f4 p1 = mat_vec(worldViewInv, {o.x, o.y, o.z, 1.0f});
    const f3 o2 = {o.x + 100.0f * dir.x, o.y + 100.0f * dir.y, o.z + 100.0f * dir.z};
    const f4 p2 = mat_vec(worldViewInv, {o2.x, o2.y, o2.z, 1.0f});
    dir = normalize({p2.x - p1.x, p2.y - p1.y, p2.z - p1.z});
    const float r[8] = {p1.x, p1.y, p1.z, 0.0f, dir.x, dir.y, dir.z, 3.402823466e+38f};
    std::memcpy(rays + 8 * t, r, sizeof(r));
    if (packedXY) packedXY[t] = XY;
  }
}

}  // extern "C"
