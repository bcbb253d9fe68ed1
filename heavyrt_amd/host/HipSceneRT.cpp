// HipSceneRT.cpp -- ISceneObject on the MI355X engine (host C++, g++).
//
// The reference registers its engines behind CreateSceneRT by name
// (core/FactoryRT.cpp:29-78) and each engine keeps geometry, instances and a
// TLAS on the host (e.g. core/BVH2CommonRT_host.cpp).  HipSceneRT keeps the
// same members' meaning but forwards everything to the C-ABI of libhrt.so:
// AddGeom_Triangles3f builds the LBVH on the device, CommitScene freezes the
// instances and picks the node layout, RayQuery_NearestHit / RayQuery_AnyHit
// answer one ray synchronously.  Renderers that can batch use the two
// *_Batch extensions (one device round trip for many rays).
//
// Errors: the reference prints and returns a default for unsupported calls
// (e.g. BVH2CommonRT.cpp:281-292, BVH2CommonRT_host.cpp:100-103); HipSceneRT
// prints the engine's error message the same way and returns the same
// defaults (geometry id 0xFFFFFFFF, a miss).
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "heavyrt/CrossRT.h"
#include "hrt.h"

namespace {

// One engine context per process (device 0), made on first use.
hrt_ctx* shared_context() {
  static std::mutex mu;
  static hrt_ctx* ctx = nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!ctx && hrt_init(0, &ctx) != HRT_OK) {
    std::printf("[HipSceneRT]: hrt_init failed: %s\n", hrt_last_error());
    ctx = nullptr;
  }
  return ctx;
}

void report(const char* where, hrt_status st) {
  if (st != HRT_OK) std::printf("[HipSceneRT::%s]: %s\n", where, hrt_last_error());
}

struct HipSceneRT : public ISceneObject {
  HipSceneRT(const char* name, int layout, const char* buildName)
      : m_name(name), m_build(buildName ? buildName : ""), m_layout(layout) {
    hrt_ctx* ctx = shared_context();
    if (ctx) report("create", hrt_scene_create(ctx, &m_scene));
  }
  ~HipSceneRT() override {
    if (m_scene) hrt_scene_destroy(m_scene);
  }

  const char* Name() const override { return m_name.c_str(); }
  const char* BuildName() const override { return m_build.c_str(); }

  void ClearGeom() override {
    if (m_scene) report("ClearGeom", hrt_scene_clear_geom(m_scene));
    m_geomBoxes.clear();
    m_instCount = 0;
    m_identity.clear();
  }

  uint32_t AddGeom_Triangles3f(const float* a_vpos3f, size_t a_vertNumber, const uint32_t* a_triIndices,
                               size_t a_indNumber, BuildQuality a_qualityLevel, size_t vByteStride) override {
    uint32_t id = 0xFFFFFFFFu;
    if (!m_scene) return id;
    hrt_status st = hrt_add_triangles(m_scene, a_vpos3f, a_vertNumber, vByteStride, a_triIndices, a_indNumber,
                                      int(a_qualityLevel), &id);
    report("AddGeom_Triangles3f", st);
    if (st != HRT_OK) return 0xFFFFFFFFu;
    // geometry box over the vertices, as BVH2CommonRT_host.cpp:78-86 keeps it
    LiteMath::Box4f box;
    const size_t stride = (vByteStride ? vByteStride : 12) / 4;
    for (size_t i = 0; i < a_vertNumber; i++)
      box.include(LiteMath::float4(a_vpos3f[i * stride], a_vpos3f[i * stride + 1], a_vpos3f[i * stride + 2], 1.0f));
    m_geomBoxes.push_back(box.boxMin);
    m_geomBoxes.push_back(box.boxMax);
    return id;
  }

  // CrossRT.h:97-105: the geometry keeps its id and may not grow; BUILD_REFIT
  // keeps the hierarchy and recomputes the boxes, other qualities rebuild.
  // The scene is traced again after the next CommitScene.
  void UpdateGeom_Triangles3f(uint32_t a_geomId, const float* a_vpos3f, size_t a_vertNumber,
                              const uint32_t* a_triIndices, size_t a_indNumber, BuildQuality a_qualityLevel,
                              size_t vByteStride) override {
    if (!m_scene) return;
    hrt_status st = hrt_update_triangles(m_scene, a_geomId, a_vpos3f, a_vertNumber, vByteStride, a_triIndices,
                                         a_indNumber, int(a_qualityLevel));
    report("UpdateGeom_Triangles3f", st);
    if (st != HRT_OK || 2 * size_t(a_geomId) + 1 >= m_geomBoxes.size()) return;
    LiteMath::Box4f box;
    const size_t stride = (vByteStride ? vByteStride : 12) / 4;
    for (size_t i = 0; i < a_vertNumber; i++)
      box.include(LiteMath::float4(a_vpos3f[i * stride], a_vpos3f[i * stride + 1], a_vpos3f[i * stride + 2], 1.0f));
    m_geomBoxes[2 * a_geomId] = box.boxMin;
    m_geomBoxes[2 * a_geomId + 1] = box.boxMax;
  }

  void ClearScene() override {
    if (m_scene) report("ClearScene", hrt_scene_clear_instances(m_scene));
    m_instCount = 0;
    m_identity.clear();
  }

  void CommitScene(BuildQuality a_qualityLevel) override {
    if (m_scene) report("CommitScene", hrt_commit(m_scene, int(a_qualityLevel), m_layout));
    ResetStats();
  }

  uint32_t AddInstance(uint32_t a_geomId, const LiteMath::float4x4& a_matrix) override {
    uint32_t id = 0xFFFFFFFFu;
    if (!m_scene) return id;
    hrt_status st = hrt_add_instance(m_scene, a_geomId, a_matrix.data(), &id);
    report("AddInstance", st);
    if (st == HRT_OK) {
      m_instCount++;
      m_identity.push_back(is_identity(a_matrix));
    }
    return st == HRT_OK ? id : 0xFFFFFFFFu;
  }

  // CrossRT.h:125-130: a new matrix for an existing instance; traced after
  // the next CommitScene.
  void UpdateInstance(uint32_t a_instanceId, const LiteMath::float4x4& a_matrix) override {
    if (!m_scene) return;
    const hrt_status st = hrt_update_instance(m_scene, a_instanceId, a_matrix.data());
    report("UpdateInstance", st);
    if (st == HRT_OK && a_instanceId < m_identity.size()) m_identity[a_instanceId] = is_identity(a_matrix);
  }

  CRT_Hit RayQuery_NearestHit(LiteMath::float4 posAndNear, LiteMath::float4 dirAndFar) override {
    CRT_Hit hit = miss(dirAndFar.w);
    if (!m_scene) return hit;
    const float p[4] = {posAndNear.x, posAndNear.y, posAndNear.z, posAndNear.w};
    const float d[4] = {dirAndFar.x, dirAndFar.y, dirAndFar.z, dirAndFar.w};
    hrt_hit h;
    hrt_status st = hrt_query_nearest_one(m_scene, p, d, &h);
    report("RayQuery_NearestHit", st);
    if (st == HRT_OK) std::memcpy(&hit, &h, sizeof(hit));
    return hit;
  }

  bool RayQuery_AnyHit(LiteMath::float4 posAndNear, LiteMath::float4 dirAndFar) override {
    // BVH4HalfRT.cpp:327-331: negate tFar and ask for the nearest hit
    dirAndFar.w = -dirAndFar.w;
    CRT_Hit h = RayQuery_NearestHit(posAndNear, dirAndFar);
    return h.geomId != 0xFFFFFFFFu;
  }

  // -- extensions: many rays per call ------------------------------------
  void RayQuery_NearestHitBatch(const LiteMath::float4* posAndNear, const LiteMath::float4* dirAndFar, size_t n,
                                CRT_Hit* out) {
    std::vector<float> rays = pack(posAndNear, dirAndFar, n);
    std::vector<uint32_t> st(2 * n);
    hrt_status s = m_scene ? hrt_trace_nearest_host(m_scene, rays.data(), n, reinterpret_cast<hrt_hit*>(out),
                                                    st.data())
                           : HRT_ERR_INVALID_ARG;
    report("RayQuery_NearestHitBatch", s);
    if (s != HRT_OK) {
      for (size_t i = 0; i < n; i++) out[i] = miss(dirAndFar[i].w);
      return;
    }
    for (size_t i = 0; i < n; i++) {
      m_pairs += st[2 * i];
      m_tris += st[2 * i + 1];
    }
    m_rays += n;
  }

  void RayQuery_AnyHitBatch(const LiteMath::float4* posAndNear, const LiteMath::float4* dirAndFar, size_t n,
                            bool* out) {
    std::vector<float> rays = pack(posAndNear, dirAndFar, n);
    std::vector<uint8_t> res(n, 0);
    hrt_status s = m_scene ? hrt_trace_any_host(m_scene, rays.data(), n, res.data()) : HRT_ERR_INVALID_ARG;
    report("RayQuery_AnyHitBatch", s);
    for (size_t i = 0; i < n; i++) out[i] = (s == HRT_OK) && res[i] != 0;
  }

  // Averages over the rays of the batched nearest-hit calls, in the
  // reference's units (FactoryRT.cpp:82-100): NC counts 2 nodes per pair
  // (BVH2CommonRT.cpp:69), BLB 32 B per node and 60 B per triangle (:70,:125).
  MetricStats GetStats() override {
    MetricStats s{};
    const double rays = m_rays ? double(m_rays) : 1.0;
    // nearest hits walk BVH4 nodes (4 per visit) only for a BVH4 or PLOC
    // layout on a scene of one identity instance; every other scene walks
    // BVH2 pairs
    const bool quads = (m_layout == HRT_LAYOUT_BVH4 || m_layout == HRT_LAYOUT_BVH4_FP16 ||
                        m_layout == HRT_LAYOUT_PLOC || m_layout == HRT_LAYOUT_PLOC_FP16) &&
                       m_identity.size() == 1 && m_identity[0];
    const double nodes = (quads ? 4.0 : 2.0) * double(m_pairs);
    s.avgNC = float(nodes / rays);
    s.avgTC = float(double(m_tris) / rays);
    s.avgBLB = float((32.0 * nodes + 60.0 * double(m_tris)) / rays);
    return s;
  }
  void ResetStats() override { m_pairs = m_tris = m_rays = 0; }

  uint32_t GetGeomNum() const override { return m_scene ? hrt_geom_count(m_scene) : 0u; }
  uint32_t GetInstNum() const override { return m_scene ? hrt_instance_count(m_scene) : 0u; }
  const LiteMath::float4* GetGeomBoxes() const override { return m_geomBoxes.empty() ? nullptr : m_geomBoxes.data(); }

 private:
  static CRT_Hit miss(float tFar) {
    CRT_Hit h;
    h.t = tFar;
    h.primId = h.instId = h.geomId = 0xFFFFFFFFu;
    h.coords[0] = h.coords[1] = h.coords[2] = h.coords[3] = 0.0f;
    return h;
  }
  static std::vector<float> pack(const LiteMath::float4* p, const LiteMath::float4* d, size_t n) {
    std::vector<float> r(8 * n);
    for (size_t i = 0; i < n; i++) {
      std::memcpy(&r[8 * i], &p[i], 16);
      std::memcpy(&r[8 * i + 4], &d[i], 16);
    }
    return r;
  }

  std::string m_name, m_build;
  int m_layout;
  hrt_scene* m_scene = nullptr;
  std::vector<LiteMath::float4> m_geomBoxes;
  uint32_t m_instCount = 0;
  uint64_t m_pairs = 0, m_tris = 0, m_rays = 0;
  std::vector<bool> m_identity;  // per instance: the matrix is the identity

  static bool is_identity(const LiteMath::float4x4& m) {
    for (int c = 0; c < 4; c++)
      for (int r = 0; r < 4; r++)
        if (m(r, c) != (r == c ? 1.0f : 0.0f)) return false;
    return true;
  }
};

}  // namespace

ISceneObject* CreateSceneRT(const char* a_implName, const char* a_buildName, const char* a_layoutName) {
  (void)a_layoutName;
  const std::string name(a_implName ? a_implName : "");
  int layout = HRT_LAYOUT_BVH2;
  if (name.find("HipPlocHalf") != std::string::npos)
    layout = HRT_LAYOUT_PLOC_FP16;
  else if (name.find("HipPloc") != std::string::npos)
    layout = HRT_LAYOUT_PLOC;
  else if (name.find("HipBVH4Half") != std::string::npos || name.find("HipBVH4_16") != std::string::npos)
    layout = HRT_LAYOUT_BVH4_FP16;
  else if (name.find("HipBVH4") != std::string::npos)
    layout = HRT_LAYOUT_BVH4;
  else if (name.find("HipAutoHalf") != std::string::npos)
    layout = HRT_LAYOUT_AUTO_FP16;
  else if (name.find("HipAuto") != std::string::npos)
    layout = HRT_LAYOUT_AUTO;
  else if (name.find("HipBVH2") == std::string::npos && name.find("HipLBVH") == std::string::npos) {
    std::printf("[ERROR!] CreateSceneRT: unknown implementation!%s\n", name.c_str());
    std::printf("Use HipBVH2\n");
  }
  return new HipSceneRT(a_implName ? a_implName : "HipBVH2", layout, a_buildName);
}

void DeleteSceneRT(ISceneObject* a_impl) { delete a_impl; }

// Batched extensions reachable through a plain ISceneObject pointer made by
// CreateSceneRT (renderer harnesses call these instead of one query per ray).
void HipSceneRT_NearestHitBatch(ISceneObject* s, const LiteMath::float4* p, const LiteMath::float4* d, size_t n,
                                CRT_Hit* out) {
  static_cast<HipSceneRT*>(s)->RayQuery_NearestHitBatch(p, d, n, out);
}
void HipSceneRT_AnyHitBatch(ISceneObject* s, const LiteMath::float4* p, const LiteMath::float4* d, size_t n,
                            bool* out) {
  static_cast<HipSceneRT*>(s)->RayQuery_AnyHitBatch(p, d, n, out);
}
