// HipSceneRT.h -- extensions of the engine's ISceneObject (no counterpart in
// the reference interface): answer many rays in one call.  `s` must come from
// CreateSceneRT of libhrt_scene.so.  Rays and results are host memory; the
// batch travels to the device once, which is what a renderer loop of
// RayQuery_NearestHit / RayQuery_AnyHit calls (eye_ray.cpp:57, rtao.cpp:86,
// 176, 183 in the reference's examples) turns into.
#pragma once

#include <cstddef>

#include "CrossRT.h"

void HipSceneRT_NearestHitBatch(ISceneObject* s, const LiteMath::float4* posAndNear,
                                const LiteMath::float4* dirAndFar, size_t n, CRT_Hit* out);
void HipSceneRT_AnyHitBatch(ISceneObject* s, const LiteMath::float4* posAndNear, const LiteMath::float4* dirAndFar,
                            size_t n, bool* out);
