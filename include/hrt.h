/* hrt.h -- C-ABI of the MI355X ray-query engine (libhrt.so).
 *
 * This is the drop-in boundary for HeavyRT's ray-query path.  In the
 * reference the boundary is the C++ interface ISceneObject plus the factory
 * CreateSceneRT (core/CrossRT.h:64-196, core/FactoryRT.cpp:29-80); every
 * entry point below names the reference member it replaces.  The C++ adapter
 * in heavyrt_amd/host/HipSceneRT.cpp implements ISceneObject on top of these
 * calls, so callers written against the reference interface run unchanged.
 *
 * Conventions
 *   - Plain C types only; every call returns hrt_status (HRT_OK == 0) and
 *     leaves a thread-local message for hrt_last_error() on failure.  No C++
 *     exception crosses this boundary.
 *   - Host input buffers (vertices, indices, matrices) are copied during the
 *     call; the caller keeps ownership (BVH2CommonRT_host.cpp:76-83).
 *   - Ray and hit buffers of the batched trace calls are DEVICE pointers
 *     (HBM), and the calls are asynchronous on the given HIP stream
 *     (a hipStream_t passed as void*; NULL = the context's stream).
 *   - Scene mutation (clear/add/commit) is exclusive; concurrent trace calls
 *     on a committed scene are allowed.
 *   - Ray record: 8 floats = posAndNear (x, y, z, tNear), dirAndFar
 *     (x, y, z, tFar), the argument pair of RayQuery_NearestHit
 *     (CrossRT.h:143).  tFar <= 0 requests an any-hit query with
 *     tFar' = -tFar that stops at the first hit (BVH4HalfRT.cpp:260-262);
 *     which hit it reports depends on the walk, so on the tree a layout
 *     walks (hit or miss does not).
 *   - primId in every hit is the ORIGINAL triangle index of the geometry
 *     (the REMAP_PRIM_ID rule, BVH2CommonRT.cpp:270-276, and BruteForceRT.cpp:255).
 */
#ifndef HRT_H_
#define HRT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hrt_status {
  HRT_OK = 0,
  HRT_ERR_INVALID_ARG = 1,   /* null pointer, bad id, bad size */
  HRT_ERR_HIP = 2,           /* a HIP runtime call failed */
  HRT_ERR_OUT_OF_MEMORY = 3, /* device allocation failed */
  HRT_ERR_NOT_COMMITTED = 4, /* trace before CommitScene */
  HRT_ERR_UNSUPPORTED = 5,   /* valid request this build does not implement */
  HRT_ERR_LEAF_OVERFLOW = 6  /* > 127 triangles share one Morton code */
} hrt_status;

typedef struct hrt_ctx hrt_ctx;
typedef struct hrt_scene hrt_scene;

/* Hit record, layout-identical to CRT_Hit (CrossRT.h:20-27): t, primId,
 * instId, geomId, coords[4] with coords[0] = u, coords[1] = v.  On a miss the
 * ids are 0xFFFFFFFF and t = tFar. */
typedef struct hrt_hit {
  float t;
  uint32_t primId;
  uint32_t instId;
  uint32_t geomId;
  float coords[4];
} hrt_hit;

/* Compact nearest-hit record (8 bytes): t and primId only. */
typedef struct hrt_hit_compact {
  float t;
  uint32_t primId;
} hrt_hit_compact;

enum { HRT_HIT_FULL = 0, HRT_HIT_COMPACT = 1 };
/* Acceleration-structure layout chosen at CommitScene:
 *   BVH2       node pairs of the LBVH; nearest hit in BVH2CommonRT's order
 *              (BVH2CommonRT.cpp:49-141), first strictly closer hit wins.
 *   BVH4       the same tree collapsed to 4-wide nodes with FP32 boxes,
 *              BVH4HalfRT's visiting order (BVH4HalfRT.cpp:73-256); nearest
 *              hit by (t, primId) so the answer does not depend on the order.
 *   BVH4_FP16  as BVH4 with FP16 boxes rounded outward (cbvh_fat.cpp:447-497);
 *              requires every coordinate within +-65504.
 *   AUTO       nearest hits on BVH2 (reference order), any-hit on BVH4.
 *   AUTO_FP16  nearest hits on BVH2 (reference order); any-hit on FP16 BVH4
 *              nodes of a second tree over the same triangles: leaves of an
 *              LBVH on a cubic Morton grid (every axis scaled by the largest
 *              extent), the hierarchy above them by agglomerative clustering;
 *              it visits fewer nodes.  Any-hit answers do not depend on the
 *              tree.  If a cubic-grid cell holds more than 127 triangles,
 *              any-hit uses the reference tree's FP16 BVH4 instead.
 *   PLOC       nearest hits and any-hit both on the FP32 BVH4 of that second
 *              tree.  Nearest hits by (t, primId), so they equal BVH4's; on
 *              4 M incoherent rays in a 4 M-triangle soup it visits a third
 *              fewer nodes.  Falls back to the reference tree's BVH4 as
 *              AUTO_FP16 does.
 *   PLOC_FP16  as PLOC with FP16 boxes (the second tree's FP16 BVH4). */
enum {
  HRT_LAYOUT_BVH2 = 0,
  HRT_LAYOUT_BVH4 = 1,
  HRT_LAYOUT_BVH4_FP16 = 2,
  HRT_LAYOUT_AUTO = 3,
  HRT_LAYOUT_AUTO_FP16 = 4,
  HRT_LAYOUT_PLOC = 5,
  HRT_LAYOUT_PLOC_FP16 = 6
};
enum { HRT_BUILD_LOW = 0, HRT_BUILD_MEDIUM = 1, HRT_BUILD_HIGH = 2, HRT_BUILD_REFIT = 3 }; /* BuildQuality, CrossRT.h:9-15 */

/* Ray/triangle test of a scene's traces (hrt_scene_set_triangle_test):
 *   MT          the reference's Moller-Trumbore with its 1e-6 barycentric
 *               slack (BVH2CommonRT.cpp:26-37); bit-exact parity mode.
 *   WATERTIGHT  the watertight FP32 test of Woop, Benthin and Wald (JCGT
 *               2(1), 2013): no slack, and no ray passes between two
 *               triangles that share an edge or a vertex (an edge value of
 *               exactly zero counts for both).  t, u and v keep their
 *               meaning.  No counterpart in the reference. */
enum { HRT_TRI_MT = 0, HRT_TRI_WATERTIGHT = 1 };

/* Debug / parity dumps (hrt_debug_dump `what`).  Available for geometries
 * added after hrt_scene_set_keep_build_arrays(scene, 1). */
enum {
  HRT_DUMP_SIZES = 0,        /* uint32[5]: triangles, leaves, max leaf size, pairs, BVH4 nodes;
                                a 6th word (if the buffer holds 24 bytes): 1 when the hierarchy
                                above the leaves is the agglomerative one (any-hit tree) */
  HRT_DUMP_CODES_INPUT = 1,  /* uint32[ntris]: Morton code per input triangle */
  HRT_DUMP_CODES_SORTED = 2, /* uint32[ntris]: codes after the (code, id) sort */
  HRT_DUMP_IDS_SORTED = 3,   /* uint32[ntris]: triangle ids after the sort */
  HRT_DUMP_RUN_START = 4,    /* uint32[leaves]: first sorted position of each unique code */
  HRT_DUMP_UNIQUE_CODES = 5, /* uint32[leaves] */
  HRT_DUMP_LR_NODES = 6,     /* uint32[(2*leaves-1)*8]: nodes in the reference's Karras
                                left/right numbering (cbvh.h:20-26; lbvh.cpp:438-439) */
  HRT_DUMP_PERM = 7,         /* uint32[ntris]: leaf-order position -> triangle id */
  HRT_DUMP_SCENE_BOX = 8,    /* float[12]: box min xyzw, box max xyzw, Morton scale xyzw */
  HRT_DUMP_QUADS32 = 9,      /* uint32[nquads*32]: BVH4 nodes, FP32 boxes (after a BVH4 commit) */
  HRT_DUMP_QUADS16 = 10,     /* uint32[nquads*16]: BVH4 nodes, FP16 boxes */
  HRT_DUMP_PAIRS = 11,       /* uint32[pairs*16]: the BVH2 node pairs as traversed */
  HRT_DUMP_ANY_TREE = 0x100  /* OR'ed into `what`: the same array of the any-hit tree (AUTO_FP16, PLOC layouts) */
};

/* Context: one device, one HIP stream.  No counterpart in the reference
 * (its engine is CPU-only); it owns what every scene shares. */
hrt_status hrt_init(int device, hrt_ctx** out);
/* Destroy every scene created on ctx before shutting it down. */
void hrt_shutdown(hrt_ctx* ctx);
/* The context's stream as a hipStream_t. */
void* hrt_ctx_stream(hrt_ctx* ctx);

/* Replaces `new X` in CreateSceneRT (FactoryRT.cpp:29-78) / DeleteSceneRT (:80). */
hrt_status hrt_scene_create(hrt_ctx* ctx, hrt_scene** out);
/* Destroy the servers that hold a scene first: a scene still held by a
 * server is kept (with a message on stderr) until the process ends. */
void hrt_scene_destroy(hrt_scene* scene);
/* Keep builder intermediates for hrt_debug_dump (parity hooks). */
hrt_status hrt_scene_set_keep_build_arrays(hrt_scene* scene, int keep);

/* ISceneObject::ClearGeom (CrossRT.h:82; BVH2CommonRT_host.cpp:20-43). */
hrt_status hrt_scene_clear_geom(hrt_scene* scene);
/* ISceneObject::ClearScene (CrossRT.h:110; BVH2CommonRT_host.cpp:105-116). */
hrt_status hrt_scene_clear_instances(hrt_scene* scene);

/* ISceneObject::AddGeom_Triangles3f (CrossRT.h:94; BVH2CommonRT_host.cpp:66-98):
 * copies nVert vertices (stride vStrideBytes, 0 = packed float3) and nIdx
 * indices, builds the LBVH on the device and returns the geometry id. */
hrt_status hrt_add_triangles(hrt_scene* scene, const float* vpos, size_t nVert, size_t vStrideBytes,
                             const uint32_t* idx, size_t nIdx, int quality, uint32_t* outGeomId);

/* ISceneObject::AddInstance (CrossRT.h:123; BVH2CommonRT_host.cpp:131-164);
 * colMajor16 is a column-major float4x4.  Fails with HRT_ERR_INVALID_ARG for
 * a singular matrix (the reference would trace with a non-finite inverse). */
hrt_status hrt_add_instance(hrt_scene* scene, uint32_t geomId, const float colMajor16[16], uint32_t* outInstId);

/* ISceneObject::UpdateGeom_Triangles3f (CrossRT.h:97-105; a stub in every
 * reference engine, e.g. BVH2CommonRT_host.cpp:100-103).  The geometry keeps
 * its id and may not grow (at most the vertex and triangle counts it was
 * added with).  quality == HRT_BUILD_REFIT keeps the hierarchy and the leaf
 * order and recomputes every box (the triangle count must not change; idx
 * may be null to keep the indices); any other quality rebuilds the LBVH.
 * The scene must be committed again before it is traced. */
hrt_status hrt_update_triangles(hrt_scene* scene, uint32_t geomId, const float* vpos, size_t nVert,
                                size_t vStrideBytes, const uint32_t* idx, size_t nIdx, int quality);

/* Warm start: a committed scene (every geometry with its trees -- LBVH node
 * pairs, leaf-order triangles, BVH4 nodes, the any-hit tree -- the
 * instances and the layout) as one byte buffer.  Call with out = NULL to get
 * the size in *size, then with a buffer of at least that size.  No
 * counterpart in the reference (its engines rebuild from geometry; the
 * reference lists serialised BVHs for serving warm start as future work). */
hrt_status hrt_scene_serialize(hrt_scene* scene, void* out, size_t capacity, size_t* size);

/* A committed scene from hrt_scene_serialize's bytes on this context's
 * device: the arrays are uploaded, nothing but an instanced scene's top level
 * is rebuilt, and traces give the saved scene's answers bit for bit.  Fails
 * with HRT_ERR_INVALID_ARG for a buffer that is not a snapshot, is truncated
 * or whose hash does not match, and with HRT_ERR_UNSUPPORTED for an unknown
 * snapshot version. */
hrt_status hrt_scene_deserialize(hrt_ctx* ctx, const void* data, size_t size, hrt_scene** out);

/* ISceneObject::UpdateInstance (CrossRT.h:125-130; a stub in the reference,
 * BVH2CommonRT_host.cpp:166-170): a new column-major matrix for an existing
 * instance; the scene must be committed again. */
hrt_status hrt_update_instance(hrt_scene* scene, uint32_t instId, const float colMajor16[16]);

/* The ray/triangle test (HRT_TRI_*) of every trace of the scene from its
 * next commit on: hrt_trace_*, the host batches, the per-ray query and the
 * AO, eye and RTAO passes.  The watertight test needs a scene with one
 * identity instance (commit fails with HRT_ERR_UNSUPPORTED otherwise); the
 * server and batches trace with the reference's test.  No counterpart in the
 * reference (its only test is BVH2CommonRT.cpp:26-37). */
hrt_status hrt_scene_set_triangle_test(hrt_scene* scene, int test);

/* ISceneObject::CommitScene (CrossRT.h:115; BVH2CommonRT_host.cpp:118-129).
 * A scene with exactly one instance whose matrix is the identity is traced on
 * its geometry's tree in `layout`.  Any other instance set gets a top-level
 * LBVH over the instance boxes (lbvh.cpp:829-917, every instance kept in its
 * leaf) and is traced two-level on the BVH2 trees for every layout
 * (BVH2CommonRT.cpp:143-279): the ray is moved into each instance's object
 * space with the inverse matrix, and on equal t the smaller instance id wins. */
hrt_status hrt_commit(hrt_scene* scene, int quality, int layout);

/* Batched ISceneObject::RayQuery_NearestHit (CrossRT.h:143; BVH2CommonRT.cpp:228-279):
 * n rays (32 B each, device) -> n hits (device, hrt_hit or hrt_hit_compact). */
hrt_status hrt_trace_nearest(hrt_scene* scene, const float* rays, size_t n, void* hits, int hitFormat,
                             void* stream);

/* Batched ISceneObject::RayQuery_AnyHit (CrossRT.h:151; BVH4HalfRT.cpp:327-331):
 * out[i] = 1 if any triangle is hit in (tNear, |tFar|). */
hrt_status hrt_trace_any(hrt_scene* scene, const float* rays, size_t n, uint8_t* out, void* stream);

/* hrt_trace_any with no allocation of any kind (no hipMalloc, no pool, no
 * allocating library call): the any-hit kernels' queue counter and stack
 * spill live in a static device array of 96 MiB + 256 B that each GPU the
 * library runs on holds until the process ends (its address is looked up
 * by hrt_init).  No counterpart in the reference: for callers that cannot
 * allocate -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation.  Scenes with one identity instance only
 * (HRT_ERR_UNSUPPORTED otherwise); any n up to 2^31-1 (the launch takes at
 * most as many workgroups as the array holds stacks for).  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU race. */
hrt_status hrt_trace_any_noalloc(hrt_scene* scene, const float* rays, size_t n, uint8_t* out, void* stream);

/* Same as hrt_trace_nearest, and also writes per-ray traversal counters
 * (uint32 pairs visited, uint32 triangles tested) -- the reference's NC/2 and
 * TC metrics (CrossRT.h:172-175, BVH2CommonRT.cpp:69,124). */
hrt_status hrt_trace_nearest_stats(hrt_scene* scene, const float* rays, size_t n, void* hits, int hitFormat,
                                   uint32_t* stats, void* stream);
hrt_status hrt_trace_any_stats(hrt_scene* scene, const float* rays, size_t n, uint8_t* out, uint32_t* stats,
                               void* stream);

/* Host-memory batches (synchronous): rays in host memory, results written to
 * host memory; staged through device buffers owned by the scene.  stats
 * (host, nullable) receives the counters of hrt_trace_nearest_stats.  For
 * callers that hold rays on the host, e.g. a renderer loop of
 * RayQuery_NearestHit / RayQuery_AnyHit calls turned into one batch. */
hrt_status hrt_trace_nearest_host(hrt_scene* scene, const float* rays, size_t n, hrt_hit* hits, uint32_t* stats);
hrt_status hrt_trace_any_host(hrt_scene* scene, const float* rays, size_t n, uint8_t* out);

/* One synchronous host query: the per-ray path behind
 * ISceneObject::RayQuery_NearestHit for callers that cannot batch. */
hrt_status hrt_query_nearest_one(hrt_scene* scene, const float posNear[4], const float dirFar[4], hrt_hit* out);

/* Device-side AO ray generation for the RTAO renderer (RTAO::kernel_InitAORay,
 * examples/02_rtao/rtao.cpp:147-165; MapSampleToCosineDistribution and
 * OffsRayPos, examples/render_common.h:56-112): for each of nPixels primary
 * rays and its compact hit, writes raysPerHit cosine-weighted rays about the
 * hit triangle's viewer-facing geometric normal with tNear = 0, tFar = tMax,
 * PCG32-seeded by `seed` and the ray index.  Missed pixels get empty rays.
 * All pointers are device pointers; asynchronous on `stream`. */
hrt_status hrt_rtao_make_rays(hrt_scene* scene, const float* primaryRays, const void* primaryHitsCompact,
                              size_t nPixels, uint32_t raysPerHit, float tMax, uint64_t seed, float* outRays,
                              void* stream);

/* The AO pass of the RTAO renderer in one call: the same rays as
 * hrt_rtao_make_rays (same arguments), traced as hrt_trace_any traces them,
 * one byte per ray in outHits (1 = occluded), ray k = pixel * raysPerHit + s.
 * With BVH4 nodes (every layout but HRT_LAYOUT_BVH2) the
 * traversal makes each ray in registers from a 32-byte per-pixel frame, so no
 * ray touches HBM; other layouts write the rays to a temporary buffer first.
 * nPixels * raysPerHit must stay below 2^32.  No counterpart in the reference
 * (its renderer calls RayQuery_AnyHit per ray). */
hrt_status hrt_rtao_trace(hrt_scene* scene, const float* primaryRays, const void* primaryHitsCompact,
                          size_t nPixels, uint32_t raysPerHit, float tMax, uint64_t seed, uint8_t* outHits,
                          void* stream);

/* Device-side eye rays of the 01_eye_rays renderer (EyeRayCaster::
 * kernel_PackXY + kernel_InitEyeRay, examples/01_eye_rays/eye_ray.cpp:26-47,
 * 114-138; EyeRayDirNormalized, transform_ray3f,
 * examples/render_common.h:6-36): one ray per pixel from the inverse
 * projection and inverse world-view matrices (column-major), pixel centres,
 * tNear = 0, tFar = FLT_MAX.  Ray k belongs to the pixel the reference's
 * SuperBlockIndex2DOpt order gives thread k (8x8 superblocks of 4x4 blocks),
 * so width and height must be multiples of 8.  outPackedXY (nullable) gets
 * y << 16 | x per ray, the reference's m_packedXY.  Device pointers;
 * asynchronous on `stream`. */
hrt_status hrt_eye_make_rays(hrt_ctx* ctx, const float projInvColMajor16[16], const float worldViewInvColMajor16[16],
                             uint32_t width, uint32_t height, float* outRays, uint32_t* outPackedXY, void* stream);

/* The 01_eye_rays frame on the device (EyeRayCaster::Render /
 * CastRaySingleBlock, eye_ray_host.cpp:329-345, with kernel_RayTrace,
 * eye_ray.cpp:50-75): eye rays as above (made and traced in an order of
 * its own rather than the superblock order, which only the image's speed
 * sees),
 * nearest hits against the committed
 * scene, and outColor[y * width + x] = 0 on a miss, else
 * palette[primId % paletteSize] (paletteSize 1..64, host array).  outColor
 * is a device buffer of width * height words; asynchronous on `stream`. */
hrt_status hrt_eye_render(hrt_scene* scene, const float projInvColMajor16[16], const float worldViewInvColMajor16[16],
                          uint32_t width, uint32_t height, const uint32_t* palette, uint32_t paletteSize,
                          uint32_t* outColor, void* stream);

/* The 02_rtao renderer's frame on the device (RTAO::CalcAO for every pixel,
 * examples/02_rtao/rtao.cpp:46-72): eye rays as hrt_eye_make_rays makes them,
 * nearest hits as hrt_trace_nearest finds them, raysPerHit AO rays per hit as
 * hrt_rtao_trace traces them, then kernel_AO2Color: outColor[y * width + x]
 * (a device image of width * height words) is 0 where the eye ray missed,
 * else the occluded fraction of its AO rays raised to `power`, clamped to
 * [0, 1] and truncated to a gray level, with alpha 255.  Width and height
 * must be multiples of 8; the scene must hold one identity instance. */
hrt_status hrt_rtao_render(hrt_scene* scene, const float projInvColMajor16[16], const float worldViewInvColMajor16[16],
                           uint32_t width, uint32_t height, uint32_t raysPerHit, float tMax, uint64_t seed, float power,
                           uint32_t* outColor, void* stream);

/* ISceneObject::GetGeomNum / GetInstNum (CrossRT.h:163-164). */
uint32_t hrt_geom_count(const hrt_scene* scene);
uint32_t hrt_instance_count(const hrt_scene* scene);

/* Parity hooks: copies one of the HRT_DUMP_* arrays of geometry geomId to
 * host memory; `bytes` is the capacity of host_out.  geomId == HRT_TLAS_ID
 * selects the top-level tree of a committed instanced scene (its "triangles"
 * are the instances). */
#define HRT_TLAS_ID 0xFFFFFFFFu
hrt_status hrt_debug_dump(hrt_scene* scene, uint32_t geomId, int what, void* host_out, size_t bytes);

/* Calibration kernel for issue-rate measurements (no reference counterpart):
 * eight waves per SIMD run chains of FP32 min / max / multiply / select with
 * no memory traffic, so under rocprofv3 its SQ_INSTS_VALU per GRBM_GUI_ACTIVE
 * cycle is the VALU issue rate such code reaches; a traversal kernel's rate
 * divided by it is that kernel's issue fraction.  Asynchronous on stream. */
hrt_status hrt_issue_probe(hrt_ctx* ctx, void* stream);

/* ---- Batches: many scenes built and traced in one HIP graph launch --------
 * No single counterpart in the reference, whose scenes are built and queried
 * one at a time (CrossRT.h:64-193).  A batch is S independent one-geometry
 * scenes, each the equivalent of AddGeom_Triangles3f (BVH2CommonRT_host.cpp:
 * 66-98) + AddInstance(identity) + CommitScene, each with its own rays traced
 * as RayQuery_NearestHit (BVH2CommonRT.cpp:228-279).  Per scene the node
 * pairs, the leaf permutation and the hits are bit-identical to a solo
 * hrt_add_triangles + hrt_trace_nearest.  Sizes are fixed at creation; the
 * caller fills the device input buffers (hrt_batch_buffers) and may refill
 * them between runs. */
typedef struct hrt_batch hrt_batch;
/* nVerts/nTris/nRays: per-scene counts (every scene >= 1 triangle). */
hrt_status hrt_batch_create(hrt_ctx* ctx, uint32_t nScenes, const uint32_t* nVerts, const uint32_t* nTris,
                            const uint32_t* nRays, hrt_batch** out);
void hrt_batch_destroy(hrt_batch* batch);
/* Device buffers, scenes concatenated in order: float4 vertices (w ignored),
 * scene-local triangle indices (3 per triangle), rays (8 floats each) and
 * compact hits (t, primId with scene-local original triangle ids). */
hrt_status hrt_batch_buffers(hrt_batch* batch, float** verts4, uint32_t** idx, float** rays, uint32_t** hits);
/* Calls that launch work (run, build, trace) are asynchronous on `stream`;
 * work on different streams must be ordered by the caller.  An index beyond
 * its scene's vertex count is clamped to the last vertex. */
/* The whole batch: every build, then every trace, as one graph launch
 * (recorded on first use and after the trace candidate changes). */
hrt_status hrt_batch_run(hrt_batch* batch, void* stream);
/* The two halves on their own, as plain launches. */
hrt_status hrt_batch_build(hrt_batch* batch, void* stream);
hrt_status hrt_batch_trace(hrt_batch* batch, void* stream);
/* Trace-kernel candidates: waves per workgroup, LDS stack entries per lane
 * (0 = one ray per lane without refill), the refill threshold (idle lanes)
 * and the number of ray indices a wave claims from the queue at once. */
int hrt_batch_candidate_count(void);
hrt_status hrt_batch_candidate_info(int i, int* wavesPerGroup, int* ldsStackEntries, int* refillAt, int* claim);
hrt_status hrt_batch_set_candidate(hrt_batch* batch, int i);
int hrt_batch_get_candidate(const hrt_batch* batch);
/* Builds if needed, times every candidate `reps` times on the batch's own
 * rays (best of reps), selects the fastest; msPerCandidate (nullable) gets
 * hrt_batch_candidate_count() timings.  Synchronous. */
hrt_status hrt_batch_autotune(hrt_batch* batch, int reps, int* chosen, float* msPerCandidate);
/* Synchronous check of the last build: HRT_OK when every scene's tree was
 * stored; HRT_ERR_LEAF_OVERFLOW (with *badScene, nullable, the first such
 * scene) when more than 127 triangles of a scene share one Morton code --
 * that scene's hits are not valid, as hrt_add_triangles would have refused
 * it; HRT_ERR_UNSUPPORTED when a device-side check of the build failed. */
hrt_status hrt_batch_status(hrt_batch* batch, uint32_t* badScene);
/* Per-scene parity hooks after a build: HRT_DUMP_SIZES (5 words: triangles,
 * leaves, longest run, pairs, first pair), HRT_DUMP_PAIRS, HRT_DUMP_PERM. */
hrt_status hrt_batch_dump(hrt_batch* batch, uint32_t scene, int what, void* host_out, size_t bytes);

/* ---- Serving: resident scenes, continuous batching ------------------------
 * No counterpart in the reference, which answers one query per virtual call
 * on the caller's thread (ISceneObject::RayQuery_NearestHit / _AnyHit,
 * CrossRT.h:143-151).  A server keeps committed scenes resident and takes
 * requests (a ray batch for one scene, nearest or any) from any thread.  Its
 * scheduler thread admits queued requests in arrival order while their
 * staging buffers (32 B per ray + the results, for buffers in host memory)
 * fit in the HBM budget, runs all admitted requests of a step as one trace
 * launch per kind of tree, keeps up to stepsInFlight steps on the device at
 * once, and admits new arrivals into the next step while earlier ones run.
 * Nearest-hit requests walk the scene's BVH2 in the reference order (compact
 * hits, equal to hrt_trace_nearest for the BVH2, AUTO and AUTO_FP16
 * layouts); any-hit requests walk the tree hrt_trace_any walks (the any-hit
 * tree's FP16 BVH4 for AUTO_FP16) and equal it.  Results are the same alone
 * or in any mix. */
typedef struct hrt_server hrt_server;
/* Pinned host memory for request rays and results (hipHostMalloc). */
hrt_status hrt_host_alloc(size_t bytes, void** out);
void hrt_host_free(void* p);
hrt_status hrt_server_create(hrt_ctx* ctx, uint64_t budgetBytes, uint32_t stepsInFlight, hrt_server** out);
/* Waits for every submitted request, then frees the server. */
void hrt_server_destroy(hrt_server* server);
/* A committed scene with one identity instance and the reference's triangle
 * test.  While a server holds it, the scene's mutators (clear, add, update,
 * commit, set_triangle_test) fail with HRT_ERR_INVALID_ARG. */
hrt_status hrt_server_add_scene(hrt_server* server, hrt_scene* scene, uint32_t* sceneId);
/* rays: n x 8 floats, out: n x hrt_hit_compact (nearest) or n bytes (any-hit);
 * each either pinned host memory (hrt_host_alloc; staged through the HBM
 * budget) or device memory of the server's device (read and written in
 * place), untouched by the caller until the request is done.  Fails with
 * HRT_ERR_OUT_OF_MEMORY if the request's staging alone exceeds the budget. */
hrt_status hrt_server_submit(hrt_server* server, uint32_t sceneId, int anyHit, const float* rays, size_t n,
                             void* out, uint64_t* requestId);
/* 1 = done, 0 = pending, negative = failed or unknown. */
int hrt_server_poll(hrt_server* server, uint64_t requestId);
/* Blocks until the request is done and forgets it; latencyMs = submit ->
 * results in host memory, queueMs = submit -> admission (both nullable). */
hrt_status hrt_server_wait(hrt_server* server, uint64_t requestId, double* latencyMs, double* queueMs);
/* out: completed requests, rays, steps, mean requests per step x 1000, peak
 * device bytes in flight, budget bytes. */
hrt_status hrt_server_stats(hrt_server* server, uint64_t out[6]);
/* out: rays launched so far on each kind of tree -- BVH2 node pairs, FP16
 * BVH4 nodes, FP32 BVH4 nodes.  Nearest-hit requests walk the node pairs;
 * any-hit requests walk the tree hrt_trace_any would (the any-hit tree of an
 * HRT_LAYOUT_AUTO_FP16 scene).  No reference counterpart (serving). */
hrt_status hrt_server_tree_rays(hrt_server* server, uint64_t out[3]);

/* Thread-local description of the last failure in this thread. */
const char* hrt_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HRT_H_ */
