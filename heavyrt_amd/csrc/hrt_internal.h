// hrt_internal.h -- host-side declarations shared by the engine's translation
// units (builder, traversal, C-ABI).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hrt.h"
#include "hrt_device.h"

// One device and one default-priority stream (hrt_init).
struct hrt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
};

namespace hrt {

// Error reporting shared by the C-ABI units: records the thread-local message
// returned by hrt_last_error and passes the status through.
hrt_status fail(hrt_status st, const std::string& msg);
hrt_status fail_hip(hipError_t e, const char* what);

struct Blas;
// The geometry tree of a committed scene with exactly one identity instance.
hrt_status scene_single_blas(hrt_scene* s, const Blas** out);
// The triangle test (HRT_TRI_*) of a committed scene's traces.
int scene_triangle_test(const hrt_scene* s);
// hrt_trace_any with caller scratch (TraceArgs::workspace): no allocation.
hrt_status trace_any_in_workspace(hrt_scene* s, const float* rays, size_t n, uint8_t* out, void* stream,
                                  uint32_t* workspace, size_t workspace_words);
// Looks up the allocation-free entry point's scratch on `device` (current),
// once per device, and sizes the any-hit launches it feeds.
hipError_t noalloc_prepare(int device);
// Queries (and caches) the resident grids of the any-hit kernels.
void warm_any_hit_launchers();
void warm_bvh2_any_hit_launchers();
// The device of a scene's context.
int scene_device(const hrt_scene* s);

// Device buffers of one bottom-level structure (one geometry).
struct Blas {
  uint32_t ntris = 0;
  uint32_t nverts = 0;
  uint32_t leaves = 0;        // unique Morton codes (valid after build)
  uint32_t max_run = 0;       // longest run of equal codes (leaf size)
  float box[8] = {0};         // scene box over vertices: min xyzw, max xyzw
  float zscale[4] = {0};      // Morton scale
  float4* verts = nullptr;    // float4 positions, w = 1
  uint32_t* idx = nullptr;    // input index buffer (3 * ntris)
  NodePair* pairs = nullptr;  // max(leaves - 1, 1) pairs
  TriPre* tris = nullptr;     // ntris, leaf order
  TriPre* tris_wt = nullptr;  // ntris, leaf order, absolute vertices {A, id}, B, C (watertight test; on demand)
  uint32_t* perm = nullptr;   // leaf-order position -> original triangle id
  uint32_t* parent = nullptr; // internal node -> (parent << 1) | side (Karras step)
  // kept from the build for refits (hierarchy unchanged, boxes recomputed)
  uint32_t* leaf_run = nullptr;     // first sorted position of each leaf's run (leaves)
  uint32_t* parent_leaf = nullptr;  // leaf -> (parent << 1) | side (leaves)
  uint2* node_range = nullptr;      // internal node -> leaf range (pairs)
  uint32_t* meta = nullptr;         // [0] leaves, [1] longest run (64 words)
  // BVH4 collapse of the same tree (made on demand by blas_build_bvh4)
  uint32_t nquads = 0;
  Quad32* quads32 = nullptr;
  Quad16* quads16 = nullptr;
  // parity hooks (kept only when requested)
  uint32_t* codes_in = nullptr;    // Morton code per input triangle
  uint32_t* codes_sorted = nullptr;
  uint32_t* ids_sorted = nullptr;
  uint32_t* run_start = nullptr;   // leaves
  uint32_t* ucodes = nullptr;      // leaves
  uint32_t* karras_lr = nullptr;   // 2 * (leaves - 1): children in the reference's left/right numbering
  bool keep_debug = false;
  // true when the hierarchy above the leaves is the agglomerative one
  // (blas_ploc): pairs, parent and karras_lr hold it; parent_leaf and
  // node_range still describe the Morton hierarchy and are not used
  bool clustered = false;
  // Second tree over the same triangles for any-hit queries: leaves on a
  // cubic Morton grid, hierarchy by blas_ploc (made on demand by
  // blas_build_any_tree; owned).
  Blas* any_tree = nullptr;
};

// Scratch for one build, sized for up to `cap` triangles; reused across builds.
struct BuildScratch {
  size_t cap = 0;
  size_t bytes = 0;
  void* base = nullptr;
};

hipError_t blas_alloc(Blas& b, uint32_t nverts, uint32_t ntris, bool keep_debug);
void blas_free(Blas& b);
// Builds the LBVH of b.verts / b.idx (already on the device) on `stream`;
// synchronises the stream before returning (leaves/max_run/box are host
// values afterwards).
hipError_t blas_build(Blas& b, hipStream_t stream, bool cubic = false);
// Refit after new vertex positions (and/or indices with the same triangle
// count) were written to b.verts / b.idx: keeps the hierarchy and the leaf
// order, recomputes every box bottom-up, the leaf-order triangles and the
// vertex box; drops the BVH4 nodes and the any-hit tree (made again on the
// next commit).  Synchronises the stream.
hipError_t blas_refit(Blas& b, hipStream_t stream);
// Builds b.any_tree (cubic grid, with its BVH4 nodes) unless it exists.
// Leaves b.any_tree null when some Morton cell of the cubic grid holds more
// than 127 triangles.  Synchronises the stream.
hipError_t blas_build_any_tree(Blas& b, hipStream_t stream);

// Replaces the hierarchy above b's leaves with the agglomerative one (PLOC,
// ploc.hip); leaves, leaf order and triangles stay.  *replaced is false (and
// b unchanged) when the tree would be deeper than the BVH4 traversal stack
// allows (kMaxLevels4 internal levels).  Synchronises the stream.
hipError_t blas_ploc(Blas& b, hipStream_t stream, bool* replaced);
// Collapses the built BVH2 into FP32 and FP16 BVH4 nodes (idempotent);
// synchronises the stream.
hipError_t blas_build_bvh4(Blas& b, hipStream_t stream);
// Leaf-order triangles with absolute vertices for the watertight test
// (idempotent).  Asynchronous on `stream`.
hipError_t blas_build_wt_tris(Blas& b, hipStream_t stream);

// Top-level LBVH over instance boxes.  `t` is allocated with blas_alloc(t,
// 2 * ninst, ninst, ...) and t.verts holds the world boxes (min, max per
// instance).  by_id[i] is instance i's record; the records are written to
// leaf_order in the tree's leaf order (t.perm maps back to instance ids).
// Synchronises the stream.
hipError_t tlas_build(Blas& t, const InstDev* by_id, InstDev* leaf_order, hipStream_t stream);

// Building blocks (graph-capturable: no allocation, no memset, no sync) -----
// Stable LSD radix sort (onesweep, sort.hip) of (key, value) pairs on the
// npasses (1-4) 8-bit digits starting at bit shift0, n < 2^30.  The sorted
// pairs end in (ka, va) after an even number of passes, in (kb, vb) after an
// odd one.  *err (zeroed by the caller) is set if a tile's look-back wait
// gives up, which would leave the order incomplete.
size_t radix_sort_scratch_words(uint64_t n, int npasses);
hipError_t radix_sort(uint32_t* ka, uint32_t* va, uint32_t* kb, uint32_t* vb, uint64_t n, int shift0, int npasses,
                      uint32_t* scratch, uint32_t* err, hipStream_t stream);
size_t scan_scratch_words(uint64_t n);
hipError_t scan_exclusive(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* temp, hipStream_t stream);
hipError_t fill_words(uint32_t* p, uint64_t n, uint32_t v, hipStream_t stream);

// Traversal ------------------------------------------------------------------
enum HitFormat : int { kHitFull = 0, kHitCompact = 1 };

// Workgroups of `kernel` (block threads, no dynamic LDS) resident on the
// current device at once: CUs x occupancy.  Queried once per device and
// kernel (a query per call shows up as launch latency on small batches).
uint32_t resident_grid(const void* kernel, int block);

struct TraceArgs {
  const Ray* rays;
  uint64_t n;
  void* out;      // CRT_Hit-compatible records, compact (t, prim) pairs or bytes
  int format;     // HitFormat for nearest-hit
  int anyhit;     // 1: boolean occlusion query
  uint32_t* stats;  // optional per-ray [pairs visited, triangles tested]
  uint32_t inst_id = 0;  // written into full hit records on a hit
  uint32_t geom_id = 0;
  bool watertight = false;  // WtTest on Blas::tris_wt instead of the reference's MtTest
  // caller-provided scratch for the persistent any-hit kernels (queue
  // counter + stack spill): when set, nothing is allocated and the launch
  // takes at most as many workgroups as the words hold stacks for
  uint32_t* workspace = nullptr;
  size_t workspace_words = 0;
};

// The triangle records a test policy reads.
template <class Tri>
inline const TriPre* tri_records(const Blas& b) {
  return std::is_same<Tri, WtTest>::value ? b.tris_wt : b.tris;
}
hipError_t trace_blas(const Blas& b, const TraceArgs& a, hipStream_t stream);

// BVH4 traversal; fp16 selects the FP16 nodes.  Nearest hit by (t, primId).
hipError_t trace_blas4(const Blas& b, const TraceArgs& a, bool fp16, hipStream_t stream);

// Two-level traversal: the TLAS (`tlas`, built by tlas_build) with instance
// records in leaf order, each instance's BVH2 in object space.  Full hit
// records carry the instance and geometry ids of the hit.
hipError_t trace_two_level(const Blas& tlas, const InstDev* inst_leaf_order, const TraceArgs& a, hipStream_t stream);

// Serving: one request inside a step, 48 bytes.  nodes is the tree the
// request's rays walk (BVH2 node pairs, or FP16/FP32 BVH4 nodes for the
// any-hit kernels), tris its leaf-order triangles; rays (device) and out
// (device: compact hits for nearest, bytes for any-hit) are the request's.
struct ServeSeg {
  const void* nodes;
  const TriPre* tris;
  const Ray* rays;
  void* out;
  uint32_t any;  // 1: any-hit query
  uint32_t pad[3];
};
static_assert(sizeof(ServeSeg) == 48, "ServeSeg is 48 bytes");
// The any-hit rays of nseg requests (ray offsets off[0..nseg]) over their
// BVH4 trees (FP16 or FP32 nodes for all of them), one persistent launch of
// at most `grid` workgroups; *queue must be zero, spill holds
// grid * 64 * (kStackDepth4 - kShortStack) words.
hipError_t serve_any4(bool fp16, const uint32_t* off, const ServeSeg* segs, uint32_t nseg, uint32_t nrays,
                      uint32_t grid, uint32_t* queue, uint32_t* spill, hipStream_t stream);
uint32_t serve_any4_resident_grid(bool fp16);
// The tree hrt_trace_any walks on a committed scene with one identity
// instance: kind 0 = BVH2 node pairs, 1 = FP16 BVH4 nodes, 2 = FP32 BVH4
// nodes.
struct TreeRef {
  const void* nodes = nullptr;
  const TriPre* tris = nullptr;
  int kind = 0;
};
TreeRef scene_any_hit_tree(const hrt_scene* s);
// A server holds the scene (count > 0): its mutators refuse to run.
void scene_serve_ref(hrt_scene* s, int delta);

// Ambient-occlusion rays from primary hits (compact records), spp per pixel.
hipError_t rtao_make_rays(const Blas& b, const Ray* prim_rays, const void* prim_hits_compact, uint64_t npix,
                          uint32_t spp, float tmax, uint64_t seed, Ray* out, hipStream_t stream);
// One AO frame per pixel (ao.h) for trace_ao4.
struct AoFrame;
hipError_t rtao_frames(const Blas& b, const Ray* prim_rays, const void* prim_hits_compact, uint64_t npix,
                       AoFrame* out, hipStream_t stream);
// The direction order of every AO sample of npix frames (ao.h, kAoOrderRun):
// npix * spp 16-bit offsets for trace_ao4.
hipError_t rtao_order(const AoFrame* frames, uint64_t npix, uint32_t spp, uint64_t seed,
                      uint16_t* order, hipStream_t stream);
// Any-hit of every AO sample of npix frames on b's BVH4 nodes (FP16 or FP32),
// each ray made in registers exactly as rtao_make_rays writes it, taken in the
// order rtao_order wrote (nullable: sample order); out gets one byte per
// sample.  npix * spp must stay below 2^32.
hipError_t trace_ao4(const Blas& b, bool fp16, bool watertight, const AoFrame* frames, const uint16_t* order,
                     uint64_t npix, uint32_t spp, float tmax, uint64_t seed, uint8_t* out, hipStream_t stream);

// Eye rays of a pinhole camera, one per pixel, in SuperBlockIndex2DOpt order
// (width and height multiples of 8); packed_xy (nullable) gets y << 16 | x.
struct EyeCamera {
  struct Mat {
    float m[16];  // column-major
  };
  Mat proj_inv, world_view_inv;
};
constexpr uint32_t kEyePaletteMax = 64;
struct EyePalette {
  uint32_t c[kEyePaletteMax];
  uint32_t n;
};
// Eye rays in the reference's superblock order, or in XCD stripes (stripes:
// the order the nearest-hit kernel traces fastest, eye.hip).
hipError_t eye_make_rays(const EyeCamera& cam, uint32_t width, uint32_t height, Ray* out, uint32_t* packed_xy,
                         hipStream_t stream, bool stripes = false);
// Pixel colours from compact hits of those rays (row-major image); stripes
// as the rays were made.
hipError_t eye_colors(const void* hits_compact, uint32_t width, uint32_t height, const EyePalette& pal,
                      uint32_t* out_color, hipStream_t stream, bool stripes = false);
// The 02_rtao frame's colours from compact primary hits (superblock order)
// and one occlusion byte per AO ray (spp per pixel).
hipError_t ao_colors(const void* hits_compact, const uint8_t* occluded, uint32_t width, uint32_t height,
                     uint32_t spp, float power, uint32_t* out_color, hipStream_t stream);

}  // namespace hrt
