// capi.hip -- implementation of the C-ABI declared in include/hrt.h.
//
// Scene model (mirrors BVH2CommonRT, core/BVH2CommonRT.h:70-85): a list of
// geometries, each with its own device BLAS, and a list of instances
// (geometry id + column-major matrix).  CommitScene freezes the instance
// list.  A scene whose only instance has the identity matrix is traced
// directly on its geometry's tree (BVH2 or BVH4 by layout); every other
// instance set gets a top-level LBVH over the instance boxes at commit and
// is traced with the two-level BVH2 kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/hrt.h"
#include "hrt_internal.h"

namespace hrt {

thread_local std::string g_last_error;

hrt_status fail(hrt_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}

hrt_status fail_hip(hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory) return fail(HRT_ERR_OUT_OF_MEMORY, std::string(what) + ": out of device memory");
  return fail(HRT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace hrt

namespace {

using hrt::fail;
using hrt::fail_hip;

struct Instance {
  uint32_t geom;
  float fwd[16];    // column-major
  float inv[12];    // rows 0..2 of the inverse
  bool identity;
};

// 4x4 inverse by the Laplace expansion in 2x2 minors (D. Eberly, "The Laplace
// Expansion Theorem: Computing the Determinants and Inverses of Matrices",
// Geometric Tools), in double, each entry rounded to float once.  The
// reference inverts in float with LiteMath's inverse4x4; evaluating in double
// gives the correctly rounded inverse for the matrices that matter here
// (rotations, scales, translations).  Writes rows 0..2; false when singular.
bool inverse_rows(const float* m, float* out12) {
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
  const double cof[12] = {
      a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3,   -a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3,
      a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3,   -a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3,
      -a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1,  a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1,
      -a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1,  a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1,
      a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0,   -a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0,
      a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0,   -a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0};
  for (int k = 0; k < 12; k++) {
    out12[k] = float(cof[k] / det);
    if (!std::isfinite(out12[k])) return false;
  }
  return true;
}

// World box of an instance: the box of the eight corners of the geometry's
// vertex box moved by the forward matrix (BVH2CommonRT_host.cpp:133-157),
// each product summed left to right as LiteMath's float4x4 * float4.
void instance_box(const float* gbox8, const float* m, float* out8) {
  const float* a = gbox8;
  const float* b = gbox8 + 4;
  const float corners[8][3] = {{a[0], a[1], a[2]}, {b[0], a[1], a[2]}, {a[0], b[1], a[2]}, {a[0], a[1], b[2]},
                               {b[0], b[1], a[2]}, {b[0], a[1], b[2]}, {a[0], b[1], b[2]}, {b[0], b[1], b[2]}};
  for (int k = 0; k < 8; k++) {
    const float* c = corners[k];
    float p[4];
    for (int r = 0; r < 4; r++) p[r] = ((m[r] * c[0] + m[4 + r] * c[1]) + m[8 + r] * c[2]) + m[12 + r] * 1.0f;
    for (int r = 0; r < 4; r++) {
      if (k == 0 || p[r] < out8[r]) out8[r] = p[r];
      if (k == 0 || out8[4 + r] < p[r]) out8[4 + r] = p[r];
    }
  }
}

// Float4 copy of caller vertices with w = 1 (BVH2CommonRT_host.cpp:79-84).
std::vector<float> repack_vertices(const float* vpos, size_t nVert, size_t vStrideBytes) {
  const size_t stride = vStrideBytes / 4;
  std::vector<float> v4(4 * nVert);
  for (size_t i = 0; i < nVert; i++) {
    v4[4 * i + 0] = vpos[i * stride + 0];
    v4[4 * i + 1] = vpos[i * stride + 1];
    v4[4 * i + 2] = vpos[i * stride + 2];
    v4[4 * i + 3] = 1.0f;
  }
  return v4;
}

bool is_identity(const float* m) {
  for (int c = 0; c < 4; c++)
    for (int r = 0; r < 4; r++)
      if (m[4 * c + r] != ((c == r) ? 1.0f : 0.0f)) return false;
  return true;
}

}  // namespace

struct hrt_scene {
  hrt_ctx* ctx = nullptr;
  std::vector<hrt::Blas> geoms;
  std::vector<Instance> instances;
  bool committed = false;
  bool keep_debug = false;
  int layout = HRT_LAYOUT_BVH2;
  int tri_test = HRT_TRI_MT;  // takes effect at the next commit
  int committed_tri_test = HRT_TRI_MT;
  // servers holding the scene (hrt_server_add_scene): its trees must not change
  std::atomic<int> served{0};
  // instanced scenes: top-level tree over the instance boxes and the
  // instance records (by instance id, then in the tree's leaf order)
  bool two_level = false;
  hrt::Blas tlas;
  hrt::InstDev* inst_dev = nullptr;
  // staging for the host-memory batch calls (grown on demand, guarded by one_mu)
  std::mutex one_mu;
  void* host_dev = nullptr;
  size_t host_dev_bytes = 0;
  // hrt_query_nearest_one: concurrent calls are traced together (below)
  struct OneQueue {
    struct Waiter {
      float ray[8];
      hrt_hit* out;
      hrt_status* status;
      std::string* message;
    };
    std::mutex mu;
    std::condition_variable cv;
    bool running = false;      // a caller is tracing a batch
    uint64_t open_gen = 1;     // the batch new calls join
    uint64_t done_gen = 0;     // the last batch that finished
    std::vector<Waiter> open;  // calls of the open batch
    // pinned host staging the trace kernel reads and writes in place
    float* rays = nullptr;
    hrt_hit* hits = nullptr;
    size_t cap = 0;
    hipStream_t stream = nullptr;
  } one;
};

namespace {

// Mutators refuse a scene that a server holds: the server walks its trees.
hrt_status check_mutable(const hrt_scene* s) {
  if (s->served.load() > 0)
    return fail(HRT_ERR_INVALID_ARG, "the scene is resident in a server (destroy the server before changing it)");
  return HRT_OK;
}

void free_tlas(hrt_scene* s) {
  if (s->tlas.pairs || s->inst_dev) {
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
  }
  hrt::blas_free(s->tlas);
  if (s->inst_dev) (void)hipFree(s->inst_dev);
  s->inst_dev = nullptr;
  s->two_level = false;
}

hrt_status build_tlas(hrt_scene* s) {
  const size_t n = s->instances.size();
  std::vector<float> boxes(8 * n);
  std::vector<hrt::InstDev> recs(n);
  for (size_t i = 0; i < n; i++) {
    const Instance& in = s->instances[i];
    const hrt::Blas& g = s->geoms[in.geom];
    instance_box(g.box, in.fwd, &boxes[8 * i]);
    hrt::InstDev& r = recs[i];
    std::memset(&r, 0, sizeof(r));
    for (int k = 0; k < 3; k++)
      r.inv_row[k] = make_float4(in.inv[4 * k + 0], in.inv[4 * k + 1], in.inv[4 * k + 2], in.inv[4 * k + 3]);
    r.pairs = g.pairs;
    r.tris = g.tris;
    r.geom = in.geom;
    r.inst = uint32_t(i);
  }
  hipStream_t hs = s->ctx->stream;
  hipError_t e = hrt::blas_alloc(s->tlas, uint32_t(2 * n), uint32_t(n), s->keep_debug);
  if (e == hipSuccess) e = hipMalloc(&s->inst_dev, 2 * n * sizeof(hrt::InstDev));
  if (e == hipSuccess) e = hipMemcpyAsync(s->tlas.verts, boxes.data(), 4 * boxes.size(), hipMemcpyHostToDevice, hs);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->inst_dev, recs.data(), n * sizeof(hrt::InstDev), hipMemcpyHostToDevice, hs);
  if (e == hipSuccess) e = hrt::tlas_build(s->tlas, s->inst_dev, s->inst_dev + n, hs);
  if (e != hipSuccess) {
    free_tlas(s);
    return fail_hip(e, "top-level build");
  }
  if (s->tlas.max_run > 127) {
    free_tlas(s);
    return fail(HRT_ERR_LEAF_OVERFLOW, "more than 127 instances share one Morton cell of the top-level tree");
  }
  s->two_level = true;
  return HRT_OK;
}

hrt_status select_blas(hrt_scene* s, const hrt::Blas** out, uint32_t* inst) {
  if (!s) return fail(HRT_ERR_INVALID_ARG, "scene is null");
  if (!s->committed) return fail(HRT_ERR_NOT_COMMITTED, "CommitScene has not been called since the last change");
  if (s->instances.size() != 1 || !s->instances[0].identity)
    return fail(HRT_ERR_UNSUPPORTED, "needs a committed scene with exactly one identity instance");
  *out = &s->geoms[s->instances[0].geom];
  *inst = 0;
  return HRT_OK;
}

hrt_status trace_common(hrt_scene* s, const float* rays, size_t n, void* out, int format, int anyhit,
                        uint32_t* stats, void* stream, uint32_t* workspace = nullptr, size_t workspace_words = 0) {
  if (!s) return fail(HRT_ERR_INVALID_ARG, "scene is null");
  if (!s->committed) return fail(HRT_ERR_NOT_COMMITTED, "CommitScene has not been called since the last change");
  const hrt::Blas* b = nullptr;
  uint32_t inst = 0;
  if (!s->two_level) {
    hrt_status st = select_blas(s, &b, &inst);
    if (st != HRT_OK) return st;
  }
  if (n == 0) return HRT_OK;
  if (!rays || !out) return fail(HRT_ERR_INVALID_ARG, "rays/out is null");
  if (n > 0x7FFFFFFFull) return fail(HRT_ERR_INVALID_ARG, "at most 2^31-1 rays per call");
  if (!anyhit && format != HRT_HIT_FULL && format != HRT_HIT_COMPACT)
    return fail(HRT_ERR_INVALID_ARG, "unknown hit format");
  hrt::TraceArgs a;
  a.rays = reinterpret_cast<const hrt::Ray*>(rays);
  a.n = n;
  a.out = out;
  a.format = format;
  a.anyhit = anyhit;
  a.stats = stats;
  a.inst_id = inst;
  a.geom_id = s->two_level ? 0u : s->instances[inst].geom;
  a.watertight = s->committed_tri_test == HRT_TRI_WATERTIGHT;
  a.workspace = workspace;
  a.workspace_words = workspace_words;
  hipStream_t hs = stream ? static_cast<hipStream_t>(stream) : s->ctx->stream;
  hipError_t e = hipSetDevice(s->ctx->device);
  if (e == hipSuccess) {
    const int lay = s->layout;
    const bool auto16 = lay == HRT_LAYOUT_AUTO_FP16;
    const bool ploc = lay == HRT_LAYOUT_PLOC || lay == HRT_LAYOUT_PLOC_FP16;
    const bool fp16 = lay == HRT_LAYOUT_BVH4_FP16 || auto16 || lay == HRT_LAYOUT_PLOC_FP16;
    // b is null for a two-level scene: every branch below it reads b
    if (s->two_level)
      e = hrt::trace_two_level(s->tlas, s->inst_dev ? s->inst_dev + s->instances.size() : nullptr, a, hs);
    else if (lay == HRT_LAYOUT_BVH2 || ((lay == HRT_LAYOUT_AUTO || auto16) && !anyhit))
      e = hrt::trace_blas(*b, a, hs);
    else if ((auto16 || ploc) && b->any_tree)
      e = hrt::trace_blas4(*b->any_tree, a, fp16, hs);  // on the second (clustered) tree
    else
      e = hrt::trace_blas4(*b, a, fp16, hs);
  }
  if (e != hipSuccess) return fail_hip(e, "trace");
  return HRT_OK;
}

}  // namespace

namespace hrt {
hrt_status scene_single_blas(hrt_scene* s, const Blas** out) {
  uint32_t inst = 0;
  return select_blas(s, out, &inst);
}
int scene_triangle_test(const hrt_scene* s) { return s->committed_tri_test; }

int scene_device(const hrt_scene* s) { return s->ctx->device; }

hrt_status trace_any_in_workspace(hrt_scene* s, const float* rays, size_t n, uint8_t* out, void* stream,
                                  uint32_t* workspace, size_t workspace_words) {
  if (!s) return fail(HRT_ERR_INVALID_ARG, "scene is null");
  if (s->committed && s->two_level)
    return fail(HRT_ERR_UNSUPPORTED, "the allocation-free any-hit trace needs a scene with one identity instance");
  return trace_common(s, rays, n, out, HRT_HIT_FULL, 1, nullptr, stream, workspace, workspace_words);
}

TreeRef scene_any_hit_tree(const hrt_scene* s) {
  // the branch trace_common takes for an any-hit batch of a single-tree scene
  const Blas& b = s->geoms[s->instances[0].geom];
  const int lay = s->layout;
  const bool auto16 = lay == HRT_LAYOUT_AUTO_FP16;
  const bool ploc = lay == HRT_LAYOUT_PLOC || lay == HRT_LAYOUT_PLOC_FP16;
  const bool fp16 = lay == HRT_LAYOUT_BVH4_FP16 || auto16 || lay == HRT_LAYOUT_PLOC_FP16;
  TreeRef t;
  t.nodes = b.pairs;
  t.tris = b.tris;
  t.kind = 0;
  if (lay == HRT_LAYOUT_BVH2) return t;
  const hrt::Blas& q = ((auto16 || ploc) && b.any_tree) ? *b.any_tree : b;
  if (fp16) {
    if (q.quads16) t = TreeRef{q.quads16, q.tris, 1};
  } else if (q.quads32) {
    t = TreeRef{q.quads32, q.tris, 2};
  }
  return t;
}

void scene_serve_ref(hrt_scene* s, int delta) { s->served += delta; }
}  // namespace hrt

extern "C" {

const char* hrt_last_error(void) { return hrt::g_last_error.c_str(); }

hrt_status hrt_init(int device, hrt_ctx** out) {
  if (!out) return fail(HRT_ERR_INVALID_ARG, "out is null");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess) return fail_hip(e, "hipGetDeviceCount");
  if (device < 0 || device >= count) return fail(HRT_ERR_INVALID_ARG, "no such device");
  if ((e = hipSetDevice(device)) != hipSuccess) return fail_hip(e, "hipSetDevice");
  hrt_ctx* c = new (std::nothrow) hrt_ctx();
  if (!c) return fail(HRT_ERR_OUT_OF_MEMORY, "host allocation failed");
  c->device = device;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
    delete c;
    return fail_hip(e, "hipStreamCreate");
  }
  // the allocation-free entry point's per-device scratch address (device current)
  if ((e = hrt::noalloc_prepare(device)) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail_hip(e, "scratch symbol");
  }
  *out = c;
  return HRT_OK;
}

void hrt_shutdown(hrt_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
}

void* hrt_ctx_stream(hrt_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

hrt_status hrt_scene_create(hrt_ctx* ctx, hrt_scene** out) {
  if (!ctx || !out) return fail(HRT_ERR_INVALID_ARG, "ctx/out is null");
  hrt_scene* s = new (std::nothrow) hrt_scene();
  if (!s) return fail(HRT_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->ctx = ctx;
  *out = s;
  return HRT_OK;
}

hrt_status hrt_scene_set_keep_build_arrays(hrt_scene* s, int keep) {
  if (!s) return fail(HRT_ERR_INVALID_ARG, "scene is null");
  s->keep_debug = keep != 0;
  return HRT_OK;
}

hrt_status hrt_scene_clear_geom(hrt_scene* s) {
  if (!s) return fail(HRT_ERR_INVALID_ARG, "scene is null");
  if (hrt_status st = check_mutable(s)) return st;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  free_tlas(s);
  for (auto& g : s->geoms) hrt::blas_free(g);
  s->geoms.clear();
  s->instances.clear();  // ClearGeom also clears the scene (BVH2CommonRT_host.cpp:42)
  s->committed = false;
  return HRT_OK;
}

void hrt_scene_destroy(hrt_scene* s) {
  if (!s) return;
  if (s->served.load() > 0) {
    // a server still walks its trees: keep everything rather than free memory in use
    std::fprintf(stderr, "hrt_scene_destroy: the scene is resident in a server; it is kept until the process ends\n");
    return;
  }
  hrt_scene_clear_geom(s);
  if (s->host_dev) (void)hipFree(s->host_dev);
  if (s->one.rays) (void)hipHostFree(s->one.rays);
  if (s->one.hits) (void)hipHostFree(s->one.hits);
  if (s->one.stream) (void)hipStreamDestroy(s->one.stream);
  delete s;
}

hrt_status hrt_scene_clear_instances(hrt_scene* s) {
  if (!s) return fail(HRT_ERR_INVALID_ARG, "scene is null");
  if (hrt_status st = check_mutable(s)) return st;
  free_tlas(s);
  s->instances.clear();
  s->committed = false;
  return HRT_OK;
}

hrt_status hrt_add_triangles(hrt_scene* s, const float* vpos, size_t nVert, size_t vStrideBytes,
                             const uint32_t* idx, size_t nIdx, int quality, uint32_t* outGeomId) {
  (void)quality;  // every quality level builds the LBVH (the reference's cbvh_lbvh preset)
  if (!s || !vpos || !idx) return fail(HRT_ERR_INVALID_ARG, "scene/vpos/idx is null");
  if (hrt_status st = check_mutable(s)) return st;
  if (vStrideBytes == 0) vStrideBytes = 12;
  if (vStrideBytes % 4 != 0 || vStrideBytes < 12) return fail(HRT_ERR_INVALID_ARG, "vertex stride must be >= 12 and a multiple of 4");
  const size_t ntris = nIdx / 3;
  if (ntris == 0) return fail(HRT_ERR_INVALID_ARG, "need at least one triangle");
  if (ntris > 0x00FFFFFFu) return fail(HRT_ERR_INVALID_ARG, "more than 16777215 triangles in one geometry (24-bit leaf start)");
  if (nVert == 0 || nVert > 0xFFFFFFFFull) return fail(HRT_ERR_INVALID_ARG, "bad vertex count");
  for (size_t i = 0; i < 3 * ntris; i++)
    if (idx[i] >= nVert) return fail(HRT_ERR_INVALID_ARG, "index out of range");

  const std::vector<float> v4 = repack_vertices(vpos, nVert, vStrideBytes);
  hipError_t e = hipSetDevice(s->ctx->device);
  if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
  hrt::Blas b;
  if ((e = hrt::blas_alloc(b, uint32_t(nVert), uint32_t(ntris), s->keep_debug)) != hipSuccess) {
    hrt::blas_free(b);
    return fail_hip(e, "AddGeom_Triangles3f allocation");
  }
  hipStream_t hs = s->ctx->stream;
  e = hipMemcpyAsync(b.verts, v4.data(), 16 * nVert, hipMemcpyHostToDevice, hs);
  if (e == hipSuccess) e = hipMemcpyAsync(b.idx, idx, 12 * ntris, hipMemcpyHostToDevice, hs);
  if (e == hipSuccess) e = hrt::blas_build(b, hs);
  if (e != hipSuccess) {
    hrt::blas_free(b);
    return fail_hip(e, "LBVH build");
  }
  if (b.max_run > 127) {
    hrt::blas_free(b);
    return fail(HRT_ERR_LEAF_OVERFLOW,
                "more than 127 triangles share one Morton code; the 7-bit leaf count (raytrace_common.h:40) cannot hold them");
  }
  s->geoms.push_back(b);
  s->committed = false;
  if (outGeomId) *outGeomId = uint32_t(s->geoms.size() - 1);
  return HRT_OK;
}

hrt_status hrt_add_instance(hrt_scene* s, uint32_t geomId, const float colMajor16[16], uint32_t* outInstId) {
  if (!s || !colMajor16) return fail(HRT_ERR_INVALID_ARG, "scene/matrix is null");
  if (hrt_status st = check_mutable(s)) return st;
  if (geomId >= s->geoms.size()) return fail(HRT_ERR_INVALID_ARG, "no such geometry");
  Instance in;
  in.geom = geomId;
  std::memcpy(in.fwd, colMajor16, sizeof(in.fwd));
  in.identity = is_identity(colMajor16);
  if (!inverse_rows(colMajor16, in.inv)) return fail(HRT_ERR_INVALID_ARG, "instance matrix is singular or not finite");
  s->instances.push_back(in);
  s->committed = false;
  if (outInstId) *outInstId = uint32_t(s->instances.size() - 1);
  return HRT_OK;
}

hrt_status hrt_update_triangles(hrt_scene* s, uint32_t geomId, const float* vpos, size_t nVert, size_t vStrideBytes,
                                const uint32_t* idx, size_t nIdx, int quality) {
  if (!s || !vpos) return fail(HRT_ERR_INVALID_ARG, "scene/vpos is null");
  if (geomId >= s->geoms.size()) return fail(HRT_ERR_INVALID_ARG, "no such geometry");
  if (hrt_status st = check_mutable(s)) return st;
  if (vStrideBytes == 0) vStrideBytes = 12;
  if (vStrideBytes % 4 != 0 || vStrideBytes < 12) return fail(HRT_ERR_INVALID_ARG, "vertex stride must be >= 12 and a multiple of 4");
  hrt::Blas& b = s->geoms[geomId];
  const size_t ntris = idx ? nIdx / 3 : b.ntris;
  if (nVert == 0 || nVert > b.nverts || ntris == 0 || ntris > b.ntris)
    return fail(HRT_ERR_INVALID_ARG, "an updated geometry may not grow (CrossRT.h:101-104) and needs a triangle");
  if (idx) {
    for (size_t i = 0; i < 3 * ntris; i++)
      if (idx[i] >= nVert) return fail(HRT_ERR_INVALID_ARG, "index out of range");
  }
  const std::vector<float> v4 = repack_vertices(vpos, nVert, vStrideBytes);
  hipError_t e = hipSetDevice(s->ctx->device);
  if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
  hipStream_t hs = s->ctx->stream;
  s->committed = false;
  free_tlas(s);
  if (quality == HRT_BUILD_REFIT) {
    if (ntris != b.ntris) return fail(HRT_ERR_INVALID_ARG, "a refit keeps the triangle count");
    if (!idx && nVert < b.nverts) {
      // the kept indices must stay inside the new vertex range: checked on the host copy
      std::vector<uint32_t> old(3 * size_t(b.ntris));
      if ((e = hipMemcpy(old.data(), b.idx, 4 * old.size(), hipMemcpyDeviceToHost)) != hipSuccess)
        return fail_hip(e, "refit");
      for (uint32_t x : old)
        if (x >= nVert) return fail(HRT_ERR_INVALID_ARG, "kept indices reach past the new vertex count");
    }
    b.nverts = uint32_t(nVert);
    e = hipMemcpyAsync(b.verts, v4.data(), 16 * nVert, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess && idx) e = hipMemcpyAsync(b.idx, idx, 12 * ntris, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hrt::blas_refit(b, hs);
    if (e != hipSuccess) return fail_hip(e, "refit");
    return HRT_OK;
  }
  // rebuild in place of the old tree (same geometry id)
  std::vector<uint32_t> kept;
  if (!idx) {
    kept.resize(3 * size_t(b.ntris));
    if ((e = hipMemcpy(kept.data(), b.idx, 4 * kept.size(), hipMemcpyDeviceToHost)) != hipSuccess)
      return fail_hip(e, "rebuild");
    for (uint32_t x : kept)
      if (x >= nVert) return fail(HRT_ERR_INVALID_ARG, "kept indices reach past the new vertex count");
    idx = kept.data();
  }
  hrt::Blas nb;
  if ((e = hrt::blas_alloc(nb, uint32_t(nVert), uint32_t(ntris), s->keep_debug)) == hipSuccess) {
    e = hipMemcpyAsync(nb.verts, v4.data(), 16 * nVert, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipMemcpyAsync(nb.idx, idx, 12 * ntris, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hrt::blas_build(nb, hs);
  }
  if (e != hipSuccess) {
    hrt::blas_free(nb);
    return fail_hip(e, "rebuild");
  }
  if (nb.max_run > 127) {
    hrt::blas_free(nb);
    return fail(HRT_ERR_LEAF_OVERFLOW, "more than 127 triangles share one Morton code");
  }
  (void)hipStreamSynchronize(hs);
  hrt::blas_free(b);
  b = nb;
  return HRT_OK;
}

hrt_status hrt_update_instance(hrt_scene* s, uint32_t instId, const float colMajor16[16]) {
  if (!s || !colMajor16) return fail(HRT_ERR_INVALID_ARG, "scene/matrix is null");
  if (hrt_status st = check_mutable(s)) return st;
  if (instId >= s->instances.size()) return fail(HRT_ERR_INVALID_ARG, "no such instance");
  Instance in = s->instances[instId];
  std::memcpy(in.fwd, colMajor16, sizeof(in.fwd));
  in.identity = is_identity(colMajor16);
  if (!inverse_rows(colMajor16, in.inv)) return fail(HRT_ERR_INVALID_ARG, "instance matrix is singular or not finite");
  s->instances[instId] = in;
  s->committed = false;
  return HRT_OK;
}

hrt_status hrt_commit(hrt_scene* s, int quality, int layout) {
  (void)quality;
  if (!s) return fail(HRT_ERR_INVALID_ARG, "scene is null");
  if (hrt_status st = check_mutable(s)) return st;
  if (layout < HRT_LAYOUT_BVH2 || layout > HRT_LAYOUT_PLOC_FP16)
    return fail(HRT_ERR_INVALID_ARG, "unknown layout");
  s->committed = false;
  free_tlas(s);
  const bool single = s->instances.size() == 1 && s->instances[0].identity;
  const bool wt = s->tri_test == HRT_TRI_WATERTIGHT;
  if (wt && !single)
    return fail(HRT_ERR_UNSUPPORTED, "the watertight triangle test traces scenes with one identity instance");
  hipError_t e = hipSetDevice(s->ctx->device);
  if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
  if (!single) {
    // instanced scenes are traced on the BVH2 trees through the top level
    if (!s->instances.empty()) {
      hrt_status st = build_tlas(s);
      if (st != HRT_OK) return st;
    }
    s->two_level = true;
  } else if (layout != HRT_LAYOUT_BVH2) {
    for (const Instance& in : s->instances) {
      hrt::Blas& b = s->geoms[in.geom];
      const bool auto16 = layout == HRT_LAYOUT_AUTO_FP16;
      const bool second = auto16 || layout == HRT_LAYOUT_PLOC || layout == HRT_LAYOUT_PLOC_FP16;
      if (layout == HRT_LAYOUT_BVH4_FP16 || auto16 || layout == HRT_LAYOUT_PLOC_FP16) {
        // FP16 boxes are only conservative inside the half range (cbvh_fat.cpp:447-448
        // maps every value beyond +-65504 to +-65504)
        for (int k = 0; k < 3; k++)
          if (!(b.box[k] >= -65504.0f && b.box[4 + k] <= 65504.0f))
            return fail(HRT_ERR_UNSUPPORTED, "FP16 nodes need every coordinate within +-65504");
      }
      if (second) {
        // the second tree on a cubic Morton grid; the reference tree's BVH4
        // only when the cubic one cannot be stored
        if ((e = hrt::blas_build_any_tree(b, s->ctx->stream)) != hipSuccess) return fail_hip(e, "any-hit tree");
        if (!b.any_tree && (e = hrt::blas_build_bvh4(b, s->ctx->stream)) != hipSuccess)
          return fail_hip(e, "BVH4 collapse");
      } else if ((e = hrt::blas_build_bvh4(b, s->ctx->stream)) != hipSuccess) {
        return fail_hip(e, "BVH4 collapse");
      }
    }
  }
  if (wt) {
    hrt::Blas& b = s->geoms[s->instances[0].geom];
    if ((e = hrt::blas_build_wt_tris(b, s->ctx->stream)) != hipSuccess ||
        (b.any_tree && (e = hrt::blas_build_wt_tris(*b.any_tree, s->ctx->stream)) != hipSuccess) ||
        (e = hipStreamSynchronize(s->ctx->stream)) != hipSuccess)
      return fail_hip(e, "watertight triangle records");
  }
  s->layout = layout;
  s->committed_tri_test = s->tri_test;
  s->committed = true;
  return HRT_OK;
}

hrt_status hrt_scene_set_triangle_test(hrt_scene* s, int test) {
  if (!s) return fail(HRT_ERR_INVALID_ARG, "scene is null");
  if (hrt_status st = check_mutable(s)) return st;
  if (test != HRT_TRI_MT && test != HRT_TRI_WATERTIGHT) return fail(HRT_ERR_INVALID_ARG, "unknown triangle test");
  s->tri_test = test;
  s->committed = false;
  return HRT_OK;
}

hrt_status hrt_trace_nearest(hrt_scene* s, const float* rays, size_t n, void* hits, int fmt, void* stream) {
  return trace_common(s, rays, n, hits, fmt, 0, nullptr, stream);
}

hrt_status hrt_trace_any(hrt_scene* s, const float* rays, size_t n, uint8_t* out, void* stream) {
  return trace_common(s, rays, n, out, HRT_HIT_FULL, 1, nullptr, stream);
}

hrt_status hrt_trace_nearest_stats(hrt_scene* s, const float* rays, size_t n, void* hits, int fmt, uint32_t* stats,
                                   void* stream) {
  if (!stats) return fail(HRT_ERR_INVALID_ARG, "stats is null");
  return trace_common(s, rays, n, hits, fmt, 0, stats, stream);
}

hrt_status hrt_trace_any_stats(hrt_scene* s, const float* rays, size_t n, uint8_t* out, uint32_t* stats,
                               void* stream) {
  if (!stats) return fail(HRT_ERR_INVALID_ARG, "stats is null");
  return trace_common(s, rays, n, out, HRT_HIT_FULL, 1, stats, stream);
}

namespace {
// Host batch: upload rays, trace, download results (and counters).
hrt_status host_batch(hrt_scene* s, const float* rays, size_t n, void* out, size_t out_bytes, int anyhit,
                      uint32_t* stats) {
  if (!s || (!rays && n) || (!out && n)) return fail(HRT_ERR_INVALID_ARG, "null argument");
  if (n == 0) return HRT_OK;
  std::lock_guard<std::mutex> lock(s->one_mu);
  const size_t ray_bytes = 32 * n, res_bytes = out_bytes * n, st_bytes = stats ? 8 * n : 0;
  const size_t need = ((ray_bytes + 255) & ~size_t(255)) + ((res_bytes + 255) & ~size_t(255)) + st_bytes;
  hipError_t e = hipSetDevice(s->ctx->device);
  if (e == hipSuccess && need > s->host_dev_bytes) {
    if (s->host_dev) (void)hipFree(s->host_dev);
    s->host_dev = nullptr;
    s->host_dev_bytes = 0;
    e = hipMalloc(&s->host_dev, need);
    if (e == hipSuccess) s->host_dev_bytes = need;
  }
  if (e != hipSuccess) return fail_hip(e, "host batch staging");
  char* base = static_cast<char*>(s->host_dev);
  float* d_rays = reinterpret_cast<float*>(base);
  void* d_out = base + ((ray_bytes + 255) & ~size_t(255));
  uint32_t* d_st = stats ? reinterpret_cast<uint32_t*>(static_cast<char*>(d_out) + ((res_bytes + 255) & ~size_t(255)))
                         : nullptr;
  hipStream_t hs = s->ctx->stream;
  if ((e = hipMemcpyAsync(d_rays, rays, ray_bytes, hipMemcpyHostToDevice, hs)) != hipSuccess)
    return fail_hip(e, "host batch upload");
  hrt_status st = trace_common(s, d_rays, n, d_out, HRT_HIT_FULL, anyhit, d_st, hs);
  if (st != HRT_OK) return st;
  if ((e = hipMemcpyAsync(out, d_out, res_bytes, hipMemcpyDeviceToHost, hs)) != hipSuccess)
    return fail_hip(e, "host batch download");
  if (stats && (e = hipMemcpyAsync(stats, d_st, st_bytes, hipMemcpyDeviceToHost, hs)) != hipSuccess)
    return fail_hip(e, "host batch download");
  if ((e = hipStreamSynchronize(hs)) != hipSuccess) return fail_hip(e, "host batch sync");
  return HRT_OK;
}
}  // namespace

hrt_status hrt_trace_nearest_host(hrt_scene* s, const float* rays, size_t n, hrt_hit* hits, uint32_t* stats) {
  return host_batch(s, rays, n, hits, sizeof(hrt_hit), 0, stats);
}

hrt_status hrt_trace_any_host(hrt_scene* s, const float* rays, size_t n, uint8_t* out) {
  return host_batch(s, rays, n, out, 1, 1, nullptr);
}

// The per-ray path.  A call joins the open batch of its scene; if no batch
// is being traced, the caller traces the open batch (its own ray and every
// ray that joined meanwhile) as one launch and hands each waiting caller its
// hit, otherwise it waits for its batch.  So concurrent per-ray calls -- a
// renderer's OpenMP loop of RayQuery_NearestHit -- share launches instead of
// queueing one launch each.  Rays and hits stay in pinned host memory that
// the kernel reads and writes in place: one launch and one wait per batch.
hrt_status hrt_query_nearest_one(hrt_scene* s, const float posNear[4], const float dirFar[4], hrt_hit* out) {
  if (!s || !posNear || !dirFar || !out) return fail(HRT_ERR_INVALID_ARG, "null argument");
  auto& q = s->one;
  hrt_status status = HRT_OK;
  std::string message;
  std::unique_lock<std::mutex> lk(q.mu);
  const uint64_t gen = q.open_gen;
  hrt_scene::OneQueue::Waiter w;
  std::memcpy(w.ray, posNear, 16);
  std::memcpy(w.ray + 4, dirFar, 16);
  w.out = out;
  w.status = &status;
  w.message = &message;
  q.open.push_back(w);
  while (q.done_gen < gen) {
    if (q.running) {
      q.cv.wait(lk);
      continue;
    }
    // lead: trace the open batch (it holds this call's ray)
    q.running = true;
    std::vector<hrt_scene::OneQueue::Waiter> batch;
    batch.swap(q.open);
    const uint64_t this_gen = q.open_gen++;
    lk.unlock();
    const size_t n = batch.size();
    hipError_t e = hipSetDevice(s->ctx->device);
    if (e == hipSuccess && !q.stream) e = hipStreamCreateWithFlags(&q.stream, hipStreamNonBlocking);
    if (e == hipSuccess && n > q.cap) {
      if (q.rays) (void)hipHostFree(q.rays);
      if (q.hits) (void)hipHostFree(q.hits);
      q.rays = nullptr;
      q.hits = nullptr;
      q.cap = 0;
      const size_t cap = std::max<size_t>(64, 2 * n);
      e = hipHostMalloc(reinterpret_cast<void**>(&q.rays), 32 * cap, hipHostMallocDefault);
      if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&q.hits), sizeof(hrt_hit) * cap, hipHostMallocDefault);
      if (e == hipSuccess) q.cap = cap;
    }
    hrt_status st = HRT_OK;
    std::string msg;
    if (e != hipSuccess) {
      st = fail_hip(e, "query staging");
    } else {
      for (size_t i = 0; i < n; i++) std::memcpy(q.rays + 8 * i, batch[i].ray, 32);
      st = trace_common(s, q.rays, n, q.hits, HRT_HIT_FULL, 0, nullptr, q.stream);
      if (st == HRT_OK && (e = hipStreamSynchronize(q.stream)) != hipSuccess) st = fail_hip(e, "query sync");
    }
    if (st != HRT_OK) msg = hrt::g_last_error;
    lk.lock();
    for (size_t i = 0; i < n; i++) {
      if (st == HRT_OK) *batch[i].out = q.hits[i];
      *batch[i].status = st;
      *batch[i].message = msg;
    }
    q.done_gen = this_gen;
    q.running = false;
    q.cv.notify_all();
  }
  lk.unlock();
  if (status != HRT_OK) return fail(status, message);
  return HRT_OK;
}

hrt_status hrt_rtao_make_rays(hrt_scene* s, const float* primaryRays, const void* primaryHitsCompact, size_t nPixels,
                              uint32_t raysPerHit, float tMax, uint64_t seed, float* outRays, void* stream) {
  const hrt::Blas* b = nullptr;
  uint32_t inst = 0;
  hrt_status st = select_blas(s, &b, &inst);
  if (st != HRT_OK) return st;
  if (nPixels == 0 || raysPerHit == 0) return HRT_OK;
  if (!primaryRays || !primaryHitsCompact || !outRays) return fail(HRT_ERR_INVALID_ARG, "null buffer");
  hipStream_t hs = stream ? static_cast<hipStream_t>(stream) : s->ctx->stream;
  hipError_t e = hipSetDevice(s->ctx->device);
  if (e == hipSuccess)
    e = hrt::rtao_make_rays(*b, reinterpret_cast<const hrt::Ray*>(primaryRays), primaryHitsCompact, nPixels,
                            raysPerHit, tMax, seed, reinterpret_cast<hrt::Ray*>(outRays), hs);
  if (e != hipSuccess) return fail_hip(e, "rtao rays");
  return HRT_OK;
}

}  // extern "C"

namespace {
// The AO pass of one frame: AO rays of every primary hit, traced for any hit
// on the tree any-hit rays use in this layout (as trace_common chooses it).
// BVH4 layouts make each ray inside the traversal from a per-pixel frame;
// BVH2 goes through a temporary ray buffer.
hipError_t rtao_pass(hrt_scene* s, const hrt::Blas* b, uint32_t inst, const hrt::Ray* prim, const void* hits,
                     uint64_t npix, uint32_t spp, float tmax, uint64_t seed, uint8_t* out, hipStream_t hs) {
  const uint64_t n = npix * spp;
  const int lay = s->layout;
  const bool auto16 = lay == HRT_LAYOUT_AUTO_FP16;
  hipError_t e = hipSuccess;
  if (lay == HRT_LAYOUT_BVH2) {
    void* rays = nullptr;
    if ((e = hipMallocAsync(&rays, sizeof(hrt::Ray) * n, hs)) != hipSuccess) return e;
    hrt::TraceArgs a;
    a.rays = static_cast<const hrt::Ray*>(rays);
    a.n = n;
    a.out = out;
    a.format = HRT_HIT_FULL;
    a.anyhit = 1;
    a.stats = nullptr;
    a.inst_id = inst;
    a.geom_id = s->instances[inst].geom;
    a.watertight = s->committed_tri_test == HRT_TRI_WATERTIGHT;
    e = hrt::rtao_make_rays(*b, prim, hits, npix, spp, tmax, seed, static_cast<hrt::Ray*>(rays), hs);
    if (e == hipSuccess) e = hrt::trace_blas(*b, a, hs);
    const hipError_t fe = hipFreeAsync(rays, hs);
    return e == hipSuccess ? fe : e;
  }
  const bool ploc = lay == HRT_LAYOUT_PLOC || lay == HRT_LAYOUT_PLOC_FP16;
  const hrt::Blas& tree = ((auto16 || ploc) && b->any_tree) ? *b->any_tree : *b;
  const bool fp16 = lay == HRT_LAYOUT_BVH4_FP16 || auto16 || lay == HRT_LAYOUT_PLOC_FP16;
  // the frames (32 B per pixel), then the samples' direction order (2 B each)
  void* frames = nullptr;
  if ((e = hipMallocAsync(&frames, 32ull * npix + 2ull * n, hs)) != hipSuccess) return e;
  hrt::AoFrame* f = static_cast<hrt::AoFrame*>(frames);
  uint16_t* order = reinterpret_cast<uint16_t*>(static_cast<char*>(frames) + 32ull * npix);
  e = hrt::rtao_frames(*b, prim, hits, npix, f, hs);
  if (e == hipSuccess) e = hrt::rtao_order(f, npix, spp, seed, order, hs);
  if (e == hipSuccess)
    e = hrt::trace_ao4(tree, fp16, s->committed_tri_test == HRT_TRI_WATERTIGHT, f, order, npix, spp, tmax, seed, out,
                       hs);
  const hipError_t fe = hipFreeAsync(frames, hs);
  return e == hipSuccess ? fe : e;
}

hrt_status rtao_counts(size_t nPixels, uint32_t raysPerHit) {
  const uint64_t n = uint64_t(nPixels) * raysPerHit;
  if (raysPerHit && (n / raysPerHit != nPixels || n > 0xFFFFFFFFull))
    return fail(HRT_ERR_INVALID_ARG, "nPixels * raysPerHit must stay below 2^32");
  return HRT_OK;
}
}  // namespace

extern "C" {

hrt_status hrt_rtao_trace(hrt_scene* s, const float* primaryRays, const void* primaryHitsCompact, size_t nPixels,
                          uint32_t raysPerHit, float tMax, uint64_t seed, uint8_t* outHits, void* stream) {
  const hrt::Blas* b = nullptr;
  uint32_t inst = 0;
  hrt_status st = select_blas(s, &b, &inst);
  if (st != HRT_OK) return st;
  if (nPixels == 0 || raysPerHit == 0) return HRT_OK;
  if (!primaryRays || !primaryHitsCompact || !outHits) return fail(HRT_ERR_INVALID_ARG, "null buffer");
  if ((st = rtao_counts(nPixels, raysPerHit)) != HRT_OK) return st;
  hipStream_t hs = stream ? static_cast<hipStream_t>(stream) : s->ctx->stream;
  hipError_t e = hipSetDevice(s->ctx->device);
  if (e == hipSuccess)
    e = rtao_pass(s, b, inst, reinterpret_cast<const hrt::Ray*>(primaryRays), primaryHitsCompact, nPixels,
                  raysPerHit, tMax, seed, outHits, hs);
  return e == hipSuccess ? HRT_OK : fail_hip(e, "rtao trace");
}

namespace {
hrt_status eye_args(const float* proj_inv, const float* world_view_inv, uint32_t width, uint32_t height,
                    hrt::EyeCamera* cam) {
  if (!proj_inv || !world_view_inv) return fail(HRT_ERR_INVALID_ARG, "camera matrix is null");
  if (width == 0 || height == 0 || (width & 7u) || (height & 7u) || width > 0xFFFFu || height > 0xFFFFu)
    return fail(HRT_ERR_INVALID_ARG, "width and height must be multiples of 8, at most 65535 (8x8 pixel tiles)");
  std::memcpy(cam->proj_inv.m, proj_inv, sizeof(cam->proj_inv.m));
  std::memcpy(cam->world_view_inv.m, world_view_inv, sizeof(cam->world_view_inv.m));
  return HRT_OK;
}
}  // namespace

hrt_status hrt_eye_make_rays(hrt_ctx* ctx, const float projInvColMajor16[16], const float worldViewInvColMajor16[16],
                             uint32_t width, uint32_t height, float* outRays, uint32_t* outPackedXY, void* stream) {
  if (!ctx || !outRays) return fail(HRT_ERR_INVALID_ARG, "null argument");
  hrt::EyeCamera cam;
  hrt_status st = eye_args(projInvColMajor16, worldViewInvColMajor16, width, height, &cam);
  if (st != HRT_OK) return st;
  hipStream_t hs = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hrt::eye_make_rays(cam, width, height, reinterpret_cast<hrt::Ray*>(outRays), outPackedXY, hs);
  if (e != hipSuccess) return fail_hip(e, "eye rays");
  return HRT_OK;
}

hrt_status hrt_eye_render(hrt_scene* s, const float projInvColMajor16[16], const float worldViewInvColMajor16[16],
                          uint32_t width, uint32_t height, const uint32_t* palette, uint32_t paletteSize,
                          uint32_t* outColor, void* stream) {
  if (!s || !outColor || !palette) return fail(HRT_ERR_INVALID_ARG, "null argument");
  if (paletteSize == 0 || paletteSize > hrt::kEyePaletteMax)
    return fail(HRT_ERR_INVALID_ARG, "the palette holds 1 to 64 colours");
  if (!s->committed) return fail(HRT_ERR_NOT_COMMITTED, "CommitScene has not been called since the last change");
  hrt::EyeCamera cam;
  hrt_status st = eye_args(projInvColMajor16, worldViewInvColMajor16, width, height, &cam);
  if (st != HRT_OK) return st;
  hrt::EyePalette pal{};
  std::memcpy(pal.c, palette, 4 * size_t(paletteSize));
  pal.n = paletteSize;
  const size_t n = size_t(width) * height;
  hipStream_t hs = stream ? static_cast<hipStream_t>(stream) : s->ctx->stream;
  hipError_t e = hipSetDevice(s->ctx->device);
  if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
  void* buf = nullptr;
  if ((e = hipMallocAsync(&buf, 40 * n, hs)) != hipSuccess) return fail_hip(e, "eye render buffers");
  float* rays = static_cast<float*>(buf);
  uint32_t* hits = reinterpret_cast<uint32_t*>(static_cast<char*>(buf) + 32 * n);
  // XCD stripes: the image is all this call owes, and stripes trace fastest
  e = hrt::eye_make_rays(cam, width, height, reinterpret_cast<hrt::Ray*>(rays), nullptr, hs, true);
  st = e == hipSuccess ? trace_common(s, rays, n, hits, HRT_HIT_COMPACT, 0, nullptr, stream) : HRT_OK;
  if (e == hipSuccess && st == HRT_OK) e = hrt::eye_colors(hits, width, height, pal, outColor, hs, true);
  const hipError_t fe = hipFreeAsync(buf, hs);
  if (st != HRT_OK) return st;
  if (e == hipSuccess) e = fe;
  if (e != hipSuccess) return fail_hip(e, "eye render");
  return HRT_OK;
}

hrt_status hrt_rtao_render(hrt_scene* s, const float projInvColMajor16[16], const float worldViewInvColMajor16[16],
                           uint32_t width, uint32_t height, uint32_t raysPerHit, float tMax, uint64_t seed, float power,
                           uint32_t* outColor, void* stream) {
  const hrt::Blas* b = nullptr;
  uint32_t inst = 0;
  hrt_status st = select_blas(s, &b, &inst);
  if (st != HRT_OK) return st;
  if (!outColor) return fail(HRT_ERR_INVALID_ARG, "null argument");
  if (raysPerHit == 0) return fail(HRT_ERR_INVALID_ARG, "raysPerHit must be at least 1");
  hrt::EyeCamera cam;
  if ((st = eye_args(projInvColMajor16, worldViewInvColMajor16, width, height, &cam)) != HRT_OK) return st;
  const size_t n = size_t(width) * height;
  if ((st = rtao_counts(n, raysPerHit)) != HRT_OK) return st;
  hipStream_t hs = stream ? static_cast<hipStream_t>(stream) : s->ctx->stream;
  hipError_t e = hipSetDevice(s->ctx->device);
  if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
  // eye rays (32 B), compact hits (8 B) and AO bytes (raysPerHit) per pixel
  void* buf = nullptr;
  if ((e = hipMallocAsync(&buf, (40 + size_t(raysPerHit)) * n, hs)) != hipSuccess)
    return fail_hip(e, "rtao render buffers");
  float* rays = static_cast<float*>(buf);
  uint32_t* hits = reinterpret_cast<uint32_t*>(static_cast<char*>(buf) + 32 * n);
  uint8_t* occ = static_cast<uint8_t*>(buf) + 40 * n;
  e = hrt::eye_make_rays(cam, width, height, reinterpret_cast<hrt::Ray*>(rays), nullptr, hs);
  st = e == hipSuccess ? trace_common(s, rays, n, hits, HRT_HIT_COMPACT, 0, nullptr, stream) : HRT_OK;
  if (e == hipSuccess && st == HRT_OK)
    e = rtao_pass(s, b, inst, reinterpret_cast<const hrt::Ray*>(rays), hits, n, raysPerHit, tMax, seed, occ, hs);
  if (e == hipSuccess && st == HRT_OK) e = hrt::ao_colors(hits, occ, width, height, raysPerHit, power, outColor, hs);
  const hipError_t fe = hipFreeAsync(buf, hs);
  if (st != HRT_OK) return st;
  if (e == hipSuccess) e = fe;
  if (e != hipSuccess) return fail_hip(e, "rtao render");
  return HRT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Scene snapshots (warm start).  Layout, little-endian:
//   "HRTSCN01", u32 version, i32 layout, u32 geometries, u32 instances,
//   per geometry a tree record, per instance u32 geometry + 16 floats
//   (column-major), u64 hash of everything before it (snap_hash).
// A tree record: u32 nverts, ntris, leaves, max_run, nquads, reserved (0),
//   clustered, has_any_tree; float box[8], zscale[4]; then the device arrays
//   at their allocated sizes (vertices, indices, pairs, triangles,
//   permutation, parents, leaf runs, leaf parents, node ranges, meta, BVH4
//   FP32 and FP16 nodes); then the any-hit tree's record if any.
// Loading uploads the arrays and commits: no LBVH, BVH4 or any-hit tree is
// rebuilt; an instanced scene's top level is (it is small).
// ---------------------------------------------------------------------------
namespace {

constexpr char kSnapMagic[8] = {'H', 'R', 'T', 'S', 'C', 'N', '0', '1'};
constexpr uint32_t kSnapVersion = 1;

// Writes into the caller's buffer (sized beforehand by snapshot_bytes); a
// write past `cap` is dropped and flagged.
struct SnapWriter {
  uint8_t* out;
  size_t cap, at = 0;
  bool overflow = false;
  bool room(size_t n) {
    if (n > cap - at) overflow = true;
    return !overflow;
  }
  void put(const void* p, size_t n) {
    if (!room(n)) return;
    std::memcpy(out + at, p, n);
    at += n;
  }
  template <class T>
  void val(T v) { put(&v, sizeof(T)); }
  hipError_t dev(const void* d, size_t n) {
    if (!room(n)) return hipSuccess;
    hipError_t e = n ? hipMemcpy(out + at, d, n, hipMemcpyDeviceToHost) : hipSuccess;
    at += n;
    return e;
  }
};

struct SnapReader {
  const uint8_t* p;
  size_t n, at = 0;
  bool get(void* out, size_t k) {
    if (k > n - at) return false;
    std::memcpy(out, p + at, k);
    at += k;
    return true;
  }
  template <class T>
  bool val(T& v) { return get(&v, sizeof(T)); }
  hipError_t dev(void* d, size_t k) {
    if (k > n - at) return hipErrorInvalidValue;
    hipError_t e = k ? hipMemcpy(d, p + at, k, hipMemcpyHostToDevice) : hipSuccess;
    at += k;
    return e;
  }
};

// FNV-1a over little-endian 64-bit words, then the tail bytes one by one
// (one multiply per 8 bytes; every step is a bijection of the state, so any
// single changed word changes the result).
uint64_t snap_hash(const uint8_t* p, size_t n) {
  constexpr uint64_t kPrime = 0x100000001b3ull;
  uint64_t h = 0xcbf29ce484222325ull;
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    std::memcpy(&w, p + i, 8);
    h = (h ^ w) * kPrime;
  }
  for (; i < n; i++) h = (h ^ p[i]) * kPrime;
  return h;
}

size_t npairs_of(const hrt::Blas& b) { return std::max<size_t>(b.ntris > 1 ? b.ntris - 1 : 1, 1); }

// Bytes of one tree record (and its any-hit tree's), as write_tree lays it out.
size_t tree_bytes(const hrt::Blas& b) {
  const size_t np = npairs_of(b), nt = b.ntris;
  size_t n = 32 + sizeof(b.box) + sizeof(b.zscale);
  n += 16ull * b.nverts + 12 * nt + sizeof(hrt::NodePair) * np + sizeof(hrt::TriPre) * nt + 4 * nt + 4 * np +
       4 * nt + 4 * nt + 8 * np + 256 + (sizeof(hrt::Quad32) + sizeof(hrt::Quad16)) * size_t(b.nquads);
  return n + (b.any_tree ? tree_bytes(*b.any_tree) : 0);
}

size_t snapshot_bytes(const hrt_scene* s) {
  size_t n = sizeof(kSnapMagic) + 16;
  for (const hrt::Blas& b : s->geoms) n += tree_bytes(b);
  return n + s->instances.size() * (4 + sizeof(float) * 16) + 8;
}

hipError_t write_tree(SnapWriter& w, const hrt::Blas& b) {
  const uint32_t head[8] = {b.nverts, b.ntris, b.leaves, b.max_run, b.nquads, 0u, b.clustered ? 1u : 0u,
                            b.any_tree ? 1u : 0u};
  w.put(head, sizeof(head));
  w.put(b.box, sizeof(b.box));
  w.put(b.zscale, sizeof(b.zscale));
  const size_t np = npairs_of(b), nt = b.ntris;
  hipError_t e = hipSuccess;
  const std::pair<const void*, size_t> arrays[] = {
      {b.verts, 16ull * b.nverts}, {b.idx, 12ull * nt}, {b.pairs, sizeof(hrt::NodePair) * np},
      {b.tris, sizeof(hrt::TriPre) * nt}, {b.perm, 4ull * nt}, {b.parent, 4ull * np},
      {b.leaf_run, 4ull * nt}, {b.parent_leaf, 4ull * nt}, {b.node_range, 8ull * np}, {b.meta, 256},
      {b.quads32, sizeof(hrt::Quad32) * b.nquads}, {b.quads16, sizeof(hrt::Quad16) * b.nquads}};
  for (const auto& a : arrays)
    if ((e = w.dev(a.first, a.second)) != hipSuccess) return e;
  return b.any_tree ? write_tree(w, *b.any_tree) : hipSuccess;
}

// Host views of one tree record's arrays inside the snapshot bytes.
struct TreeSpans {
  const uint32_t* idx;
  const hrt::NodePair* pairs;
  const hrt::TriPre* tris;
  const uint32_t* perm;
  const uint32_t* parent;
  const uint32_t* leaf_run;
  const uint32_t* parent_leaf;
  const uint2* node_range;
  const uint32_t* meta;
  const hrt::Quad32* quads32;
  const hrt::Quad16* quads16;
};

// Child references of one tree level form a tree: every internal reference
// below `count`, no node but the root (0) referenced, every other node
// referenced exactly once, every leaf inside the triangle array, and no
// internal node deeper than `max_depth` (the traversal stacks are sized for
// that depth).  refs(k, out) writes node k's `fan` child references.
template <class Refs>
const char* check_tree(uint32_t count, int fan, uint32_t ntris, uint32_t max_depth, Refs refs) {
  std::vector<uint32_t> depth(count, 0xFFFFFFFFu);
  std::vector<uint32_t> order;
  order.reserve(count);
  depth[0] = 0;
  order.push_back(0);
  uint32_t ch[4];
  for (size_t at = 0; at < order.size(); at++) {
    const uint32_t k = order[at];
    refs(k, ch);
    for (int i = 0; i < fan; i++) {
      const uint32_t r = ch[i];
      if (r & hrt::kLeafBit) {
        if (uint64_t(hrt::leaf_start(r)) + hrt::leaf_count(r) > ntris) return "a leaf reaches past the triangles";
        continue;
      }
      if (r == 0 || r >= count) return "a child reference is out of range";
      if (depth[r] != 0xFFFFFFFFu) return "a node is referenced twice";
      depth[r] = depth[k] + 1;
      if (depth[r] > max_depth) return "the tree is deeper than the traversal stack allows";
      order.push_back(r);
    }
  }
  return nullptr;
}

// Every index and reference a kernel follows stays inside its array, and the
// trees are trees of bounded depth, so a corrupted or crafted snapshot cannot
// make a kernel read or write out of bounds or loop.
const char* check_tree_record(const uint32_t head[8], const TreeSpans& t, bool any_tree) {
  const uint32_t nverts = head[0], ntris = head[1], leaves = head[2], max_run = head[3], nquads = head[4];
  if (max_run == 0 || max_run > 127) return "bad longest leaf";
  if (t.meta[0] != leaves || t.meta[1] != max_run) return "meta words disagree with the header";
  for (size_t i = 0; i < 3ull * ntris; i++)
    if (t.idx[i] >= nverts) return "an index reaches past the vertices";
  std::vector<uint8_t> seen(ntris, 0);
  for (uint32_t i = 0; i < ntris; i++) {
    const uint32_t p = t.perm[i];
    if (p >= ntris || seen[p]) return "the leaf permutation is not a permutation";
    seen[p] = 1;
    if (__builtin_bit_cast(uint32_t, t.tris[i].a.w) >= ntris) return "a triangle id is out of range";
  }
  const uint32_t np = leaves > 1 ? leaves - 1 : 1;
  for (uint32_t k = 0; k < np; k++) {
    if (k > 0 && (t.parent[k] >> 1) >= np) return "a parent reference is out of range";
    if (leaves > 1 && (t.node_range[k].x > t.node_range[k].y || t.node_range[k].y >= leaves))
      return "a node range is out of range";
  }
  for (uint32_t i = 0; i < leaves; i++)  // a one-leaf tree has no parent records
    if (t.leaf_run[i] >= ntris || (leaves > 1 && (t.parent_leaf[i] >> 1) >= np)) return "a leaf record is out of range";
  // node pairs: the BVH2 kernels keep kStackDepth pending entries, at most one per level
  const char* err = check_tree(np, 2, ntris, any_tree ? 0xFFFFFFFFu : uint32_t(hrt::kStackDepth - 1),
                               [&](uint32_t k, uint32_t* c) {
                                 c[0] = __builtin_bit_cast(uint32_t, t.pairs[k].lmin.w);
                                 c[1] = __builtin_bit_cast(uint32_t, t.pairs[k].rmin.w);
                               });
  if (err) return err;
  if (nquads) {
    // BVH4 nodes: at most three pending entries per level in kStackDepth4
    const uint32_t d4 = uint32_t(hrt::kStackDepth4 / 3 - 1);
    if ((err = check_tree(nquads, 4, ntris, d4, [&](uint32_t k, uint32_t* c) {
           c[0] = t.quads32[k].ref.x;
           c[1] = t.quads32[k].ref.y;
           c[2] = t.quads32[k].ref.z;
           c[3] = t.quads32[k].ref.w;
         })))
      return err;
    for (uint32_t k = 0; k < nquads; k++)
      for (int i = 0; i < 4; i++)
        if (t.quads16[k].c[i].w != (&t.quads32[k].ref.x)[i]) return "FP16 and FP32 nodes disagree";
  }
  return nullptr;
}

// Reads one tree record into b (allocated here).  The record is checked on
// the host before anything is uploaded.  Returns HRT_OK or an error status
// with the message set.
hrt_status read_tree(SnapReader& r, hrt::Blas& b, int depth) {
  uint32_t head[8];
  float box[8], zs[4];
  if (!r.val(head) || !r.val(box) || !r.val(zs)) return fail(HRT_ERR_INVALID_ARG, "snapshot is truncated");
  const uint32_t nverts = head[0], ntris = head[1];
  if (nverts == 0 || ntris == 0 || ntris > 0x00FFFFFFu || head[2] == 0 || head[2] > ntris || head[5] != 0 ||
      head[7] > 1 || (head[7] && depth > 0))
    return fail(HRT_ERR_INVALID_ARG, "snapshot holds an invalid tree record");
  const size_t np = std::max<size_t>(ntris > 1 ? ntris - 1 : 1, 1), nt = ntris, nq = head[4];
  const size_t sizes[] = {16ull * nverts, 12ull * nt, sizeof(hrt::NodePair) * np, sizeof(hrt::TriPre) * nt,
                          4ull * nt, 4ull * np, 4ull * nt, 4ull * nt, 8ull * np, 256,
                          sizeof(hrt::Quad32) * nq, sizeof(hrt::Quad16) * nq};
  constexpr int kArrays = sizeof(sizes) / sizeof(sizes[0]);
  const uint8_t* at[kArrays];
  size_t off = r.at;
  for (int i = 0; i < kArrays; i++) {
    if (sizes[i] > r.n - off) return fail(HRT_ERR_INVALID_ARG, "snapshot is truncated");
    at[i] = r.p + off;
    off += sizes[i];
  }
  // the arrays are read in place; records are 4-byte aligned in the buffer
  // when it is, and copied otherwise
  std::vector<uint8_t> copy;
  if ((reinterpret_cast<uintptr_t>(r.p + r.at) & 15u) != 0) {
    copy.assign(r.p + r.at, r.p + off);
    size_t o = 0;
    for (int i = 0; i < kArrays; i++) {
      at[i] = copy.data() + o;
      o += sizes[i];
    }
  }
  const TreeSpans spans{reinterpret_cast<const uint32_t*>(at[1]),      reinterpret_cast<const hrt::NodePair*>(at[2]),
                        reinterpret_cast<const hrt::TriPre*>(at[3]),   reinterpret_cast<const uint32_t*>(at[4]),
                        reinterpret_cast<const uint32_t*>(at[5]),      reinterpret_cast<const uint32_t*>(at[6]),
                        reinterpret_cast<const uint32_t*>(at[7]),      reinterpret_cast<const uint2*>(at[8]),
                        reinterpret_cast<const uint32_t*>(at[9]),      reinterpret_cast<const hrt::Quad32*>(at[10]),
                        reinterpret_cast<const hrt::Quad16*>(at[11])};
  if (const char* bad = check_tree_record(head, spans, depth > 0))
    return fail(HRT_ERR_INVALID_ARG, std::string("snapshot holds an invalid tree record: ") + bad);
  hipError_t e = hrt::blas_alloc(b, nverts, ntris, false);
  if (e != hipSuccess) return fail_hip(e, "snapshot allocation");
  b.leaves = head[2];
  b.max_run = head[3];
  b.clustered = head[6] != 0;
  std::memcpy(b.box, box, sizeof(b.box));
  std::memcpy(b.zscale, zs, sizeof(b.zscale));
  if (nq && ((e = hipMalloc(&b.quads32, sizeof(hrt::Quad32) * nq)) != hipSuccess ||
             (e = hipMalloc(&b.quads16, sizeof(hrt::Quad16) * nq)) != hipSuccess))
    return fail_hip(e, "snapshot allocation");
  b.nquads = head[4];
  void* dst[kArrays] = {b.verts, b.idx,         b.pairs,      b.tris, b.perm,    b.parent,
                        b.leaf_run, b.parent_leaf, b.node_range, b.meta, b.quads32, b.quads16};
  for (int i = 0; i < kArrays; i++)
    if (sizes[i] && (e = hipMemcpy(dst[i], at[i], sizes[i], hipMemcpyHostToDevice)) != hipSuccess)
      return fail_hip(e, "snapshot upload");
  r.at = off;
  if (head[7]) {
    b.any_tree = new (std::nothrow) hrt::Blas();
    if (!b.any_tree) return fail(HRT_ERR_OUT_OF_MEMORY, "host allocation failed");
    return read_tree(r, *b.any_tree, depth + 1);
  }
  return HRT_OK;
}

}  // namespace

extern "C" {

hrt_status hrt_scene_serialize(hrt_scene* s, void* out, size_t capacity, size_t* size) {
  if (!s || !size) return fail(HRT_ERR_INVALID_ARG, "scene/size is null");
  if (!s->committed) return fail(HRT_ERR_NOT_COMMITTED, "only a committed scene is saved");
  const size_t total = snapshot_bytes(s);
  *size = total;
  if (!out) return HRT_OK;
  if (capacity < total) return fail(HRT_ERR_INVALID_ARG, "output buffer too small (query the size first)");
  SnapWriter w{static_cast<uint8_t*>(out), total};
  w.put(kSnapMagic, sizeof(kSnapMagic));
  w.val(kSnapVersion);
  w.val(int32_t(s->layout | (s->committed_tri_test << 8)));
  w.val(uint32_t(s->geoms.size()));
  w.val(uint32_t(s->instances.size()));
  hipError_t e = hipSetDevice(s->ctx->device);
  if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
  for (size_t g = 0; e == hipSuccess && g < s->geoms.size(); g++) e = write_tree(w, s->geoms[g]);
  if (e != hipSuccess) return fail_hip(e, "snapshot download");
  for (const Instance& in : s->instances) {
    w.val(uint32_t(in.geom));
    w.put(in.fwd, sizeof(in.fwd));
  }
  if (w.overflow || w.at + 8 != total) return fail(HRT_ERR_INVALID_ARG, "internal: snapshot size mismatch");
  w.val(snap_hash(w.out, w.at));
  return HRT_OK;
}

hrt_status hrt_scene_deserialize(hrt_ctx* ctx, const void* data, size_t size, hrt_scene** out) {
  if (!ctx || !data || !out) return fail(HRT_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  const uint8_t* p = static_cast<const uint8_t*>(data);
  if (size < sizeof(kSnapMagic) + 24 || std::memcmp(p, kSnapMagic, sizeof(kSnapMagic)) != 0)
    return fail(HRT_ERR_INVALID_ARG, "not a scene snapshot");
  uint64_t want = 0;
  std::memcpy(&want, p + size - 8, 8);
  if (snap_hash(p, size - 8) != want) return fail(HRT_ERR_INVALID_ARG, "scene snapshot is corrupted (hash mismatch)");
  SnapReader r{p, size - 8};
  r.at = sizeof(kSnapMagic);
  uint32_t version = 0, ngeom = 0, ninst = 0;
  int32_t layout = 0;
  if (!r.val(version) || !r.val(layout) || !r.val(ngeom) || !r.val(ninst))
    return fail(HRT_ERR_INVALID_ARG, "snapshot is truncated");
  if (version != kSnapVersion) return fail(HRT_ERR_UNSUPPORTED, "unknown snapshot version");
  hrt_scene* s = nullptr;
  hrt_status st = hrt_scene_create(ctx, &s);
  if (st != HRT_OK) return st;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) {
    hrt_scene_destroy(s);
    return fail_hip(e, "hipSetDevice");
  }
  for (uint32_t g = 0; st == HRT_OK && g < ngeom; g++) {
    s->geoms.emplace_back();
    st = read_tree(r, s->geoms.back(), 0);
  }
  for (uint32_t i = 0; st == HRT_OK && i < ninst; i++) {
    uint32_t geom = 0;
    float m[16];
    if (!r.val(geom) || !r.val(m)) st = fail(HRT_ERR_INVALID_ARG, "snapshot is truncated");
    else st = hrt_add_instance(s, geom, m, nullptr);
  }
  if (st == HRT_OK && r.at != r.n) st = fail(HRT_ERR_INVALID_ARG, "snapshot has trailing bytes");
  if (st == HRT_OK) st = hrt_scene_set_triangle_test(s, layout >> 8);
  if (st == HRT_OK) st = hrt_commit(s, HRT_BUILD_MEDIUM, layout & 0xFF);
  if (st != HRT_OK) {
    hrt_scene_destroy(s);
    return st;
  }
  *out = s;
  return HRT_OK;
}

uint32_t hrt_geom_count(const hrt_scene* s) { return s ? uint32_t(s->geoms.size()) : 0u; }
uint32_t hrt_instance_count(const hrt_scene* s) { return s ? uint32_t(s->instances.size()) : 0u; }

hrt_status hrt_debug_dump(hrt_scene* s, uint32_t geomId, int what, void* host_out, size_t bytes) {
  if (!s || !host_out) return fail(HRT_ERR_INVALID_ARG, "scene/out is null");
  if (geomId == HRT_TLAS_ID) {
    if (!s->committed || !s->tlas.pairs) return fail(HRT_ERR_INVALID_ARG, "no top-level tree (commit an instanced scene)");
  } else if (geomId >= s->geoms.size()) {
    return fail(HRT_ERR_INVALID_ARG, "no such geometry");
  }
  const bool any_tree = (what & HRT_DUMP_ANY_TREE) != 0;
  what &= ~HRT_DUMP_ANY_TREE;
  if (any_tree && (geomId == HRT_TLAS_ID || !s->geoms[geomId].any_tree))
    return fail(HRT_ERR_INVALID_ARG, "no any-hit tree (commit with HRT_LAYOUT_AUTO_FP16 or a PLOC layout)");
  const hrt::Blas& b = (geomId == HRT_TLAS_ID) ? s->tlas : (any_tree ? *s->geoms[geomId].any_tree : s->geoms[geomId]);
  const uint32_t N = b.ntris, L = b.leaves;
  const uint32_t npairs = std::max<uint32_t>(L > 1 ? L - 1 : 1, 1);
  auto need = [&](size_t n) -> hrt_status {
    if (bytes < n) return fail(HRT_ERR_INVALID_ARG, "output buffer too small");
    return HRT_OK;
  };
  hipError_t e = hipSetDevice(s->ctx->device);
  if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
  if (what == HRT_DUMP_SIZES) {
    if (need(20)) return HRT_ERR_INVALID_ARG;
    uint32_t v[6] = {N, L, b.max_run, npairs, b.nquads, b.clustered ? 1u : 0u};
    std::memcpy(host_out, v, bytes >= 24 ? 24 : 20);
    return HRT_OK;
  }
  if (what == HRT_DUMP_SCENE_BOX) {
    if (need(48)) return HRT_ERR_INVALID_ARG;
    std::memcpy(host_out, b.box, 32);
    std::memcpy(static_cast<char*>(host_out) + 32, b.zscale, 16);
    return HRT_OK;
  }
  if (what == HRT_DUMP_QUADS32 || what == HRT_DUMP_QUADS16) {
    if (!b.quads32) return fail(HRT_ERR_INVALID_ARG, "no BVH4 nodes (commit with a BVH4 layout first)");
    const size_t n = (what == HRT_DUMP_QUADS32 ? sizeof(hrt::Quad32) : sizeof(hrt::Quad16)) * size_t(b.nquads);
    if (need(n)) return HRT_ERR_INVALID_ARG;
    e = hipMemcpy(host_out, what == HRT_DUMP_QUADS32 ? static_cast<const void*>(b.quads32) : b.quads16, n,
                  hipMemcpyDeviceToHost);
    return e == hipSuccess ? HRT_OK : fail_hip(e, "dump");
  }
  if (what == HRT_DUMP_PAIRS) {
    if (need(64ull * npairs)) return HRT_ERR_INVALID_ARG;
    e = hipMemcpy(host_out, b.pairs, 64ull * npairs, hipMemcpyDeviceToHost);
    return e == hipSuccess ? HRT_OK : fail_hip(e, "dump");
  }
  if (what == HRT_DUMP_PERM) {
    if (need(4ull * N)) return HRT_ERR_INVALID_ARG;
    e = hipMemcpy(host_out, b.perm, 4ull * N, hipMemcpyDeviceToHost);
    return e == hipSuccess ? HRT_OK : fail_hip(e, "dump");
  }
  if (!b.keep_debug) return fail(HRT_ERR_INVALID_ARG, "build arrays were not kept (hrt_scene_set_keep_build_arrays)");
  const void* src = nullptr;
  size_t n = 0;
  switch (what) {
    case HRT_DUMP_CODES_INPUT: src = b.codes_in; n = 4ull * N; break;
    case HRT_DUMP_CODES_SORTED: src = b.codes_sorted; n = 4ull * N; break;
    case HRT_DUMP_IDS_SORTED: src = b.ids_sorted; n = 4ull * N; break;
    case HRT_DUMP_RUN_START: src = b.run_start; n = 4ull * L; break;
    case HRT_DUMP_UNIQUE_CODES: src = b.ucodes; n = 4ull * L; break;
    case HRT_DUMP_LR_NODES: break;
    default: return fail(HRT_ERR_INVALID_ARG, "unknown dump");
  }
  if (src) {
    if (need(n)) return HRT_ERR_INVALID_ARG;
    e = hipMemcpy(host_out, src, n, hipMemcpyDeviceToHost);
    return e == hipSuccess ? HRT_OK : fail_hip(e, "dump");
  }
  // Reconstruct the reference's left/right node array (2L-1 nodes of 8 words)
  // from the node pairs: internal node k = union of pair k's boxes, children
  // from the Karras step; leaf m = the box its parent pair holds.
  const size_t nodes = 2ull * L - 1;
  if (need(nodes * 32)) return HRT_ERR_INVALID_ARG;
  std::vector<uint32_t> pw(16ull * npairs), lr(2ull * npairs), rs(L);
  if ((e = hipMemcpy(pw.data(), b.pairs, 64ull * npairs, hipMemcpyDeviceToHost)) != hipSuccess) return fail_hip(e, "dump");
  if ((e = hipMemcpy(rs.data(), b.run_start, 4ull * L, hipMemcpyDeviceToHost)) != hipSuccess) return fail_hip(e, "dump");
  if (L > 1 && (e = hipMemcpy(lr.data(), b.karras_lr, 8ull * (L - 1), hipMemcpyDeviceToHost)) != hipSuccess)
    return fail_hip(e, "dump");
  uint32_t* o = static_cast<uint32_t*>(host_out);
  auto fmin_ = [](uint32_t a, uint32_t c) {
    float x, y;
    std::memcpy(&x, &a, 4);
    std::memcpy(&y, &c, 4);
    float r = (y < x) ? y : x;
    uint32_t u;
    std::memcpy(&u, &r, 4);
    return u;
  };
  auto fmax_ = [](uint32_t a, uint32_t c) {
    float x, y;
    std::memcpy(&x, &a, 4);
    std::memcpy(&y, &c, 4);
    float r = (x < y) ? y : x;
    uint32_t u;
    std::memcpy(&u, &r, 4);
    return u;
  };
  if (L == 1) {
    for (int k = 0; k < 8; k++) o[k] = pw[k];
    o[3] = pw[3];
    o[7] = 0xFFFFFFFFu;
    return HRT_OK;
  }
  for (uint32_t k = 0; k + 1 < L; k++) {
    const uint32_t* p = &pw[16ull * k];
    uint32_t* nd = &o[8ull * k];
    for (int a = 0; a < 3; a++) {
      nd[a] = fmin_(p[a], p[8 + a]);
      nd[4 + a] = fmax_(p[4 + a], p[12 + a]);
    }
    nd[3] = lr[2 * k];
    nd[7] = lr[2 * k + 1];
    for (int side = 0; side < 2; side++) {
      const uint32_t child = lr[2 * k + side];
      if (child >= L - 1) {  // leaf
        const uint32_t m = child - (L - 1);
        uint32_t* lf = &o[8ull * child];
        for (int a = 0; a < 3; a++) {
          lf[a] = p[8 * side + a];
          lf[4 + a] = p[8 * side + 4 + a];
        }
        lf[3] = p[8 * side + 3];
        lf[7] = 0xFFFFFFFFu;
        (void)m;
      }
    }
  }
  return HRT_OK;
}

}  // extern "C"
