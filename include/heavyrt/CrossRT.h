// CrossRT.h -- the ray-query interface of HeavyRT, declared for the MI355X
// engine.  Types, members and signatures are the reference's
// (core/CrossRT.h:9-27 BuildQuality/CRT_Hit, :41-58 MetricStats, :64-193
// ISceneObject, :195-196 factory) so renderers written against it compile
// and run unchanged; the implementation behind CreateSceneRT is the GPU
// engine of heavyrt_amd/host/HipSceneRT.cpp over the C-ABI in hrt.h.
#pragma once

#include <cstddef>
#include <cstdint>

#include "LiteMath.h"

enum BuildQuality {
  BUILD_LOW = 0,     ///< prefer fast build
  BUILD_MEDIUM = 1,  ///< standard sweep builder
  BUILD_HIGH = 2,    ///< split BVH / early split clipping in the reference; LBVH here
  BUILD_REFIT = 3,   ///< keep the hierarchy, recompute boxes
};

/// Hit record: t, primitive / instance / geometry ids, barycentrics u = coords[0], v = coords[1].
struct CRT_Hit {
  float t;
  uint32_t primId;
  uint32_t instId;
  uint32_t geomId;
  float coords[4];
};

constexpr int TREELET_ARR_SIZE = 5;

/// Per-ray averages of the traversal counters (core/CrossRT.h:41-58).
struct MetricStats {
  float avgNC;   ///< nodes visited
  float avgLC;   ///< leaves visited
  float avgTC;   ///< triangles tested
  float avgLJC[TREELET_ARR_SIZE];
  float avgCMC[TREELET_ARR_SIZE];
  uint32_t avgWSS[TREELET_ARR_SIZE];
  float avgBLB;  ///< bus load in bytes
  float avgSOC;
  float avgSBL;
  size_t bvhTotalSize;
  size_t geomTotalSize;
};

struct ISceneObject {
  ISceneObject() {}
  virtual ~ISceneObject() {}

  virtual const char* Name() const = 0;
  virtual const char* BuildName() const { return nullptr; }

  virtual void ClearGeom() = 0;
  virtual uint32_t AddGeom_Triangles3f(const float* a_vpos3f, size_t a_vertNumber, const uint32_t* a_triIndices,
                                       size_t a_indNumber, BuildQuality a_qualityLevel = BUILD_HIGH,
                                       size_t vByteStride = sizeof(float) * 3) = 0;
  virtual void UpdateGeom_Triangles3f(uint32_t a_geomId, const float* a_vpos3f, size_t a_vertNumber,
                                      const uint32_t* a_triIndices, size_t a_indNumber,
                                      BuildQuality a_qualityLevel = BUILD_HIGH,
                                      size_t vByteStride = sizeof(float) * 3) = 0;
  virtual void ClearScene() = 0;
  virtual void CommitScene(BuildQuality a_qualityLevel = BUILD_MEDIUM) = 0;
  virtual uint32_t AddInstance(uint32_t a_geomId, const LiteMath::float4x4& a_matrix) = 0;
  virtual void UpdateInstance(uint32_t a_instanceId, const LiteMath::float4x4& a_matrix) = 0;

  virtual CRT_Hit RayQuery_NearestHit(LiteMath::float4 posAndNear, LiteMath::float4 dirAndFar) = 0;
  virtual bool RayQuery_AnyHit(LiteMath::float4 posAndNear, LiteMath::float4 dirAndFar) = 0;

  virtual MetricStats GetStats() { return MetricStats{}; }
  virtual void ResetStats() {}

  virtual uint32_t GetGeomNum() const { return 0; }
  virtual uint32_t GetInstNum() const { return 0; }
  virtual const LiteMath::float4* GetGeomBoxes() const { return nullptr; }
};

/// Factory by implementation name (core/FactoryRT.cpp:29-78).  The engine's
/// names: "HipBVH2" (LBVH node pairs, BVH2CommonRT order), "HipBVH4"
/// (BVH4, FP32 boxes), "HipBVH4Half" (BVH4, FP16 boxes), "HipAuto" (nearest
/// on BVH2, any-hit on BVH4), "HipAutoHalf" (nearest on BVH2, any-hit on the
/// FP16 BVH4 of a clustered second tree), "HipPloc" / "HipPlocHalf" (both
/// queries on the second tree's FP32 / FP16 BVH4).  Any other name prints an
/// error and returns the
/// "HipBVH2" engine, as the reference falls back to its default engine.
ISceneObject* CreateSceneRT(const char* a_implName, const char* a_buildName, const char* a_layoutName);
void DeleteSceneRT(ISceneObject* a_impl);
