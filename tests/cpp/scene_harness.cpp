// scene_harness.cpp -- drives the engine through ISceneObject the way the
// reference's renderers do (examples/01_eye_rays/eye_ray_host.cpp:283-310 for
// the build, eye_ray.cpp:57 for one RayQuery_NearestHit per pixel,
// rtao.cpp:183 for RayQuery_AnyHit), then repeats the queries through the
// batch extensions.
//
//   scene_harness <impl> <mesh.bin> <rays.bin> <out.bin> [threads]
//   mesh.bin: u32 nVert, u32 nIdx, float4 vertices[nVert], u32 indices[nIdx]
//   rays.bin: u32 n, float rays[n][8] (posAndNear, dirAndFar)
//   out.bin : CRT_Hit per-ray[n], CRT_Hit batch[n], u8 anyPerRay[n],
//             u8 anyBatch[n], u32 geomNum, u32 instNum
//   threads  > 0: also runs the per-ray loop from that many OpenMP threads
//            (eye_ray_host.cpp:327-363 issues RayQuery_NearestHit from an
//            OpenMP loop), times every call and the whole loop, checks the
//            hits against the sequential per-ray pass and prints one JSON
//            line {"per_ray_loop": ...}
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "heavyrt/CrossRT.h"
#include "heavyrt/HipSceneRT.h"

using LiteMath::float4;
using LiteMath::float4x4;

template <class T>
static bool read_exact(FILE* f, T* dst, size_t n) {
  return std::fread(dst, sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
  if (argc != 5 && argc != 6) {
    std::fprintf(stderr, "usage: %s impl mesh.bin rays.bin out.bin [threads]\n", argv[0]);
    return 2;
  }
  const int threads = argc == 6 ? std::atoi(argv[5]) : 0;
  FILE* fm = std::fopen(argv[2], "rb");
  FILE* fr = std::fopen(argv[3], "rb");
  if (!fm || !fr) return 3;
  uint32_t nv = 0, ni = 0, n = 0;
  if (!read_exact(fm, &nv, 1) || !read_exact(fm, &ni, 1)) return 3;
  std::vector<float4> verts(nv);
  std::vector<uint32_t> idx(ni);
  if (!read_exact(fm, verts.data(), nv) || !read_exact(fm, idx.data(), ni)) return 3;
  if (!read_exact(fr, &n, 1)) return 3;
  std::vector<float4> pos(n), dir(n);
  for (uint32_t i = 0; i < n; i++)
    if (!read_exact(fr, &pos[i], 1) || !read_exact(fr, &dir[i], 1)) return 3;
  std::fclose(fm);
  std::fclose(fr);

  ISceneObject* scene = CreateSceneRT(argv[1], "cbvh_lbvh", "");
  scene->ClearGeom();
  const uint32_t geomId = scene->AddGeom_Triangles3f(reinterpret_cast<const float*>(verts.data()), nv, idx.data(), ni,
                                                     BUILD_HIGH, sizeof(float4));
  scene->ClearScene();
  scene->AddInstance(geomId, float4x4());
  scene->CommitScene();

  std::vector<CRT_Hit> per_ray(n), batch(n);
  std::vector<uint8_t> any_per_ray(n), any_batch(n);
  for (uint32_t i = 0; i < n; i++) per_ray[i] = scene->RayQuery_NearestHit(pos[i], dir[i]);
  for (uint32_t i = 0; i < n; i++) any_per_ray[i] = scene->RayQuery_AnyHit(pos[i], dir[i]) ? 1 : 0;
  HipSceneRT_NearestHitBatch(scene, pos.data(), dir.data(), n, batch.data());
  std::vector<char> tmp(n);
  HipSceneRT_AnyHitBatch(scene, pos.data(), dir.data(), n, reinterpret_cast<bool*>(tmp.data()));
  for (uint32_t i = 0; i < n; i++) any_batch[i] = tmp[i] ? 1 : 0;
  if (threads > 0) {
    using clk = std::chrono::steady_clock;
    std::vector<CRT_Hit> par(n);
    std::vector<float> us(n);
    for (int pass = 0; pass < 2; pass++) {  // the first pass warms the staging up
      const auto t0 = clk::now();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64)
      for (int64_t i = 0; i < int64_t(n); i++) {
        const auto a = clk::now();
        par[i] = scene->RayQuery_NearestHit(pos[i], dir[i]);
        us[i] = std::chrono::duration<float, std::micro>(clk::now() - a).count();
      }
      const double loop_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
      if (pass == 0) continue;
      uint32_t mism = 0;
      for (uint32_t i = 0; i < n; i++) mism += std::memcmp(&par[i], &per_ray[i], sizeof(CRT_Hit)) != 0;
      std::vector<float> srt(us);
      std::sort(srt.begin(), srt.end());
      std::printf("{\"per_ray_loop\": {\"rays\": %u, \"threads\": %d, \"loop_ms\": %.3f, \"mrays_s\": %.4f, "
                  "\"call_us_p50\": %.2f, \"call_us_p99\": %.2f, \"mismatches_vs_sequential\": %u}}\n",
                  n, threads, loop_ms, n / (loop_ms * 1e3), srt[n / 2], srt[std::min<size_t>(n - 1, size_t(0.99 * n))],
                  mism);
    }
  }
  const MetricStats st = scene->GetStats();
  const uint32_t counts[2] = {scene->GetGeomNum(), scene->GetInstNum()};
  std::printf("%s: %u rays, avg nodes %.2f, avg tris %.2f, avg bytes %.1f\n", scene->Name(), n, st.avgNC, st.avgTC,
              st.avgBLB);
  DeleteSceneRT(scene);

  FILE* fo = std::fopen(argv[4], "wb");
  if (!fo) return 4;
  std::fwrite(per_ray.data(), sizeof(CRT_Hit), n, fo);
  std::fwrite(batch.data(), sizeof(CRT_Hit), n, fo);
  std::fwrite(any_per_ray.data(), 1, n, fo);
  std::fwrite(any_batch.data(), 1, n, fo);
  std::fwrite(counts, sizeof(uint32_t), 2, fo);
  std::fclose(fo);
  return 0;
}
