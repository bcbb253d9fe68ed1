// serve.hip -- a request loop over resident scenes with continuous batching.
//
// The reference answers one query per virtual call on the caller's thread
// (ISceneObject::RayQuery_NearestHit / RayQuery_AnyHit, core/CrossRT.h:143-151).
// A server keeps many committed scenes resident in HBM and takes requests --
// a batch of rays for one scene, nearest-hit or any-hit -- from any thread:
//
//   submit   the request joins a FIFO with its arrival time
//   admit    the scheduler thread forms a STEP from the head of the FIFO:
//            requests are admitted while their staging buffers fit in what
//            is left of the HBM budget (first-fit in one device arena) and
//            the step has room; a request that does not fit waits for memory
//            freed by earlier ones, the order is kept.  Requests whose rays
//            and results are already device buffers need no staging.
//   run      one step = copy the host requests' rays in, then one persistent
//            trace launch per kind of tree over all of the step's requests
//            of that kind (each ray finds its request by a binary search over
//            the step's ray offsets), then copy the host requests' results
//            out; up to `slots` steps are in flight on their own streams, so
//            requests that arrive while others are on the GPU are admitted
//            into the next step and their work overlaps on the device
//   retire   when a step's event has completed its requests are done and
//            their memory returns to the budget
//
// Each request's rays walk the tree hrt_trace_nearest / hrt_trace_any walk
// on that scene: nearest hits on the BVH2 in BVH2CommonRT's order
// (BVH2CommonRT.cpp:49-141, first strictly closer hit), any-hit on the
// scene's any-hit tree (the BVH4 of the clustered second tree for
// HRT_LAYOUT_AUTO_FP16 and the PLOC layouts, else its BVH4 nodes, else its
// BVH2), stopping at the first hit.  So a result is the same alone and in
// any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "hrt_internal.h"

namespace hrt {
namespace {

constexpr int kWaveS = 64;
constexpr int kRefillS = 16;
constexpr uint32_t kClaimS = 128;
constexpr uint32_t kMaxPerStep = 256;

// Step table in device memory: [queue counter + pad: 64 words]
// [ray offsets: kMaxPerStep + 1 words, padded] [segments].
constexpr size_t kTabOffsWords = 64;
constexpr size_t kTabSegByte = 4 * (64 + ((kMaxPerStep + 1 + 63) & ~size_t(63)));
constexpr size_t kTabBytes = kTabSegByte + sizeof(ServeSeg) * kMaxPerStep;

__device__ inline uint32_t seg_of_step(const uint32_t* __restrict__ off, uint32_t k, uint32_t i) {
  uint32_t lo = 0, hi = k;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kWaveS) void k_serve(const uint8_t* __restrict__ table, uint32_t nseg, uint32_t R,
                                                  uint32_t* __restrict__ spill) {
  __shared__ uint32_t stack_mem[kShortStack * kWaveS];
  uint32_t* queue = const_cast<uint32_t*>(reinterpret_cast<const uint32_t*>(table));
  const uint32_t* off = reinterpret_cast<const uint32_t*>(table) + kTabOffsWords;
  const ServeSeg* segs = reinterpret_cast<const ServeSeg*>(table + kTabSegByte);
  const ShortStack<kShortStack, kWaveS> stk{stack_mem + threadIdx.x, spill + blockIdx.x * kWaveS + threadIdx.x,
                                            gridDim.x * kWaveS};
  bool active = false;
  uint32_t rid = 0, loc = 0;
  const NodePair* P = nullptr;
  const TriPre* Tr = nullptr;
  void* out = nullptr;
  bool any = false, any_out = false;
  float3 o = make_float3(0.f, 0.f, 0.f), d = o, inv = o;
  float tnear = 0.f;
  HitRec h;
  h.t = 0.f;
  h.prim = kInvalid;
  h.u = h.v = 0.f;
  int top = 0;
  uint32_t cur = kInvalid;
  RayPool pool;
  while (true) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = __popcll(idle);
    if (!pool.drained() && (nidle >= kRefillS || nidle == 64)) {
      const uint64_t idx = pool.take<kClaimS>(nidle, lanes_below(idle), queue, R);
      if (!active && idx < R) {
        rid = uint32_t(idx);
        const uint32_t sg = seg_of_step(off, nseg, rid);
        const ServeSeg& S = segs[sg];
        loc = rid - off[sg];
        P = static_cast<const NodePair*>(S.nodes);
        Tr = S.tris;
        out = S.out;
        any_out = S.any != 0;
        const Ray r = S.rays[loc];
        o = make_float3(r.pos_near.x, r.pos_near.y, r.pos_near.z);
        d = make_float3(r.dir_far.x, r.dir_far.y, r.dir_far.z);
        tnear = r.pos_near.w;
        any = any_out || r.dir_far.w <= 0.0f;
        h.t = any_out ? fabsf(r.dir_far.w) : (any ? -r.dir_far.w : r.dir_far.w);
        h.prim = kInvalid;
        inv = safe_inverse(d);
        top = 0;
        cur = 0;
        active = true;
      }
    }
    if (__ballot(active) == 0) break;
    while (true) {
      bool done = false;
      if (active) {
        if ((cur & kLeafBit) == 0) {
          const float4* pp = reinterpret_cast<const float4*>(P + cur);
          const float4 lmin = pp[0], lmax = pp[1], rmin = pp[2], rmax = pp[3];
          const float2 t0 = slab(o, inv, lmin.x, lmin.y, lmin.z, lmax.x, lmax.y, lmax.z);
          const float2 t1 = slab(o, inv, rmin.x, rmin.y, rmin.z, rmax.x, rmax.y, rmax.z);
          const bool h0 = (t0.x <= t0.y) && (t0.y >= tnear) && (t0.x <= h.t);
          const bool h1 = (t1.x <= t1.y) && (t1.y >= tnear) && (t1.x <= h.t);
          const uint32_t c0 = __float_as_uint(lmin.w), c1 = __float_as_uint(rmin.w);
          if (h0 && h1) {
            const bool left_first = t0.x <= t1.x;
            cur = left_first ? c0 : c1;
            stk.push(top, left_first ? c1 : c0);
          } else if (h0 || h1) {
            cur = h0 ? c0 : c1;
          } else {
            cur = (top > 0) ? stk.pop(top) : kInvalid;
          }
        } else if (cur != kInvalid) {
          const uint32_t s = leaf_start(cur), c = leaf_count(cur);
          for (uint32_t k = s; k < s + c; k++)
            if (mt_test(o, d, tnear, Tr[k], h) && any) break;
          cur = (top > 0) ? stk.pop(top) : kInvalid;
        }
        done = (cur == kInvalid) || (any && h.prim != kInvalid);
      }
      if (done) {
        if (any_out) static_cast<uint8_t*>(out)[loc] = (h.prim != kInvalid) ? 1 : 0;
        else static_cast<uint2*>(out)[loc] = make_uint2(__float_as_uint(h.t), h.prim);
        active = false;
      }
      const uint32_t nact = __popcll(__ballot(active));
      if (nact == 0 || (!pool.drained() && 64 - nact >= kRefillS)) break;
    }
  }
}

using Clock = std::chrono::steady_clock;

}  // namespace
}  // namespace hrt

// ===========================================================================
// server object
// ===========================================================================
namespace {

using hrt::fail;
using hrt::fail_hip;

constexpr int kKinds = 3;  // trees a step launches over: BVH2 pairs, FP16 BVH4, FP32 BVH4

struct Request {
  uint64_t id = 0;
  uint32_t scene = 0;
  int any = 0;
  const float* rays = nullptr;  // pinned host or device
  void* out = nullptr;          // pinned host or device
  bool rays_dev = false, out_dev = false;
  uint32_t n = 0;
  size_t bytes = 0;             // staging bytes in the arena (host buffers only)
  size_t dev_off = 0;
  hrt::Clock::time_point t_submit, t_admit, t_done;
  int state = 0;                // 0 queued, 1 running, 2 done, 3 failed
};

struct Arena {  // first-fit over [0, size) with coalescing free list
  size_t size = 0;
  std::map<size_t, size_t> free;  // offset -> length
  void init(size_t sz) {
    size = sz;
    free.clear();
    free[0] = sz;
  }
  bool alloc(size_t n, size_t* off) {
    if (n == 0) {
      *off = 0;
      return true;
    }
    for (auto it = free.begin(); it != free.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n, o = it->first + n;
        free.erase(it);
        if (rest) free[o] = rest;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    if (n == 0) return;
    auto it = free.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free.erase(nx);
    }
    if (it != free.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free.erase(it);
      }
    }
  }
};

struct Step {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  uint8_t* tab_host = nullptr;  // pinned, kKinds tables
  uint8_t* tab_dev = nullptr;
  uint32_t* spill = nullptr;
  std::vector<uint64_t> reqs;
  bool busy = false;
};

// A resident scene as the server reads it (pointers taken at hrt_server_add_scene;
// the scene refuses changes while a server holds it).
struct ServedScene {
  hrt_scene* scene = nullptr;
  const hrt::NodePair* pairs = nullptr;
  const hrt::TriPre* tris = nullptr;
  hrt::TreeRef any;  // the tree any-hit requests walk
};

}  // namespace

struct hrt_server {
  hrt_ctx* ctx = nullptr;
  std::vector<ServedScene> scenes;
  size_t budget = 0;
  uint8_t* arena_dev = nullptr;
  Arena arena;
  std::vector<Step> steps;
  uint32_t grid[kKinds] = {1, 1, 1};
  size_t spill_words = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> queue;
  std::map<uint64_t, Request> reqs;
  uint64_t next_id = 1;
  bool stop = false;
  std::thread worker;
  // statistics
  uint64_t completed = 0, failed = 0, rays_done = 0, nsteps = 0, step_reqs = 0;
  uint64_t kind_rays[kKinds] = {0, 0, 0};  // rays launched per tree kind
  size_t in_flight = 0, peak_in_flight = 0;
  hipError_t fault = hipSuccess;
};

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

size_t result_bytes(uint32_t n, int any) { return (any ? 1ull : 8ull) * n; }

// Arena bytes of a request: staging for whichever of its buffers is in host memory.
size_t staging_bytes(const Request& r) {
  return (r.rays_dev ? 0 : align_up(32ull * r.n, 256)) + (r.out_dev ? 0 : align_up(result_bytes(r.n, r.any), 256));
}

// Launches one step with the admitted requests (lock held by the caller).
hipError_t launch_step(hrt_server* sv, Step& st) {
  using namespace hrt;
  const uint32_t k = uint32_t(st.reqs.size());
  uint32_t nk[kKinds] = {0, 0, 0}, rk[kKinds] = {0, 0, 0};
  hipError_t e = hipSuccess;
  for (int t = 0; t < kKinds; t++) std::memset(st.tab_host + t * kTabBytes, 0, 256);  // queue counters
  for (uint32_t i = 0; i < k && e == hipSuccess; i++) {
    Request& q = sv->reqs[st.reqs[i]];
    const ServedScene& sc = sv->scenes[q.scene];
    uint8_t* base = sv->arena_dev + q.dev_off;
    const Ray* d_rays = q.rays_dev ? reinterpret_cast<const Ray*>(q.rays) : reinterpret_cast<const Ray*>(base);
    void* d_out = q.out_dev ? q.out : base + (q.rays_dev ? 0 : align_up(32ull * q.n, 256));
    ServeSeg sg;
    std::memset(&sg, 0, sizeof(sg));
    int kind = 0;
    if (q.any && sc.any.kind != 0) {
      kind = sc.any.kind;
      sg.nodes = sc.any.nodes;
      sg.tris = sc.any.tris;
    } else {
      sg.nodes = sc.pairs;
      sg.tris = sc.tris;
    }
    sg.rays = d_rays;
    sg.out = d_out;
    sg.any = q.any ? 1u : 0u;
    uint8_t* tab = st.tab_host + kind * kTabBytes;
    uint32_t* off = reinterpret_cast<uint32_t*>(tab) + kTabOffsWords;
    off[nk[kind]] = rk[kind];
    reinterpret_cast<ServeSeg*>(tab + kTabSegByte)[nk[kind]] = sg;
    nk[kind]++;
    rk[kind] += q.n;
    if (!q.rays_dev) e = hipMemcpyAsync(base, q.rays, 32ull * q.n, hipMemcpyHostToDevice, st.stream);
  }
  for (int t = 0; t < kKinds; t++) (reinterpret_cast<uint32_t*>(st.tab_host + t * kTabBytes) + kTabOffsWords)[nk[t]] = rk[t];
  if (e == hipSuccess)
    e = hipMemcpyAsync(st.tab_dev, st.tab_host, kKinds * kTabBytes, hipMemcpyHostToDevice, st.stream);
  for (int t = 0; t < kKinds && e == hipSuccess; t++) {
    if (rk[t] == 0) continue;
    uint8_t* tab = st.tab_dev + t * kTabBytes;
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(sv->grid[t], (rk[t] + kWaveS - 1) / kWaveS));
    if (t == 0) {
      k_serve<<<grid, kWaveS, 0, st.stream>>>(tab, nk[t], rk[t], st.spill);
      e = hipGetLastError();
    } else {
      e = serve_any4(t == 1, reinterpret_cast<const uint32_t*>(tab) + kTabOffsWords,
                     reinterpret_cast<const ServeSeg*>(tab + kTabSegByte), nk[t], rk[t], grid,
                     reinterpret_cast<uint32_t*>(tab), st.spill, st.stream);
    }
    if (e == hipSuccess) sv->kind_rays[t] += rk[t];
  }
  for (uint32_t i = 0; i < k && e == hipSuccess; i++) {
    Request& q = sv->reqs[st.reqs[i]];
    if (q.out_dev) continue;
    uint8_t* base = sv->arena_dev + q.dev_off + (q.rays_dev ? 0 : align_up(32ull * q.n, 256));
    e = hipMemcpyAsync(q.out, base, result_bytes(q.n, q.any), hipMemcpyDeviceToHost, st.stream);
  }
  if (e == hipSuccess) e = hipEventRecord(st.done, st.stream);
  return e;
}

// Completes a step's requests (state 2 done, 3 failed) and returns their
// memory to the budget.
void retire_step(hrt_server* sv, Step& st, bool ok) {
  const auto now = hrt::Clock::now();
  for (uint64_t id : st.reqs) {
    Request& r = sv->reqs[id];
    r.state = ok ? 2 : 3;
    r.t_done = now;
    sv->arena.release(r.dev_off, r.bytes);
    sv->in_flight -= r.bytes;
    if (ok) {
      sv->completed++;
      sv->rays_done += r.n;
    } else {
      sv->failed++;
    }
  }
  st.reqs.clear();
  st.busy = false;
}

void worker_loop(hrt_server* sv) {
  (void)hipSetDevice(sv->ctx->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  while (true) {
    // retire finished steps
    bool progressed = false;
    for (Step& st : sv->steps) {
      if (!st.busy) continue;
      hipError_t q = hipEventQuery(st.done);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess && sv->fault == hipSuccess) sv->fault = q;
      retire_step(sv, st, q == hipSuccess);
      progressed = true;
    }
    if (progressed) sv->cv_done.notify_all();
    // admit into a free step slot
    if (!sv->queue.empty()) {
      for (Step& st : sv->steps) {
        if (st.busy || sv->queue.empty()) continue;
        while (!sv->queue.empty() && st.reqs.size() < hrt::kMaxPerStep) {
          Request& r = sv->reqs[sv->queue.front()];
          size_t o = 0;
          if (!sv->arena.alloc(r.bytes, &o)) break;  // keep FIFO order: wait for memory
          r.dev_off = o;
          r.state = 1;
          r.t_admit = hrt::Clock::now();
          sv->in_flight += r.bytes;
          sv->peak_in_flight = std::max(sv->peak_in_flight, sv->in_flight);
          st.reqs.push_back(r.id);
          sv->queue.pop_front();
        }
        if (st.reqs.empty()) break;
        sv->nsteps++;
        sv->step_reqs += st.reqs.size();
        st.busy = true;
        hipError_t e = launch_step(sv, st);
        if (e != hipSuccess) {
          // nothing of the step may have run: its requests fail now (an
          // event that was never recorded would read as complete)
          if (sv->fault == hipSuccess) sv->fault = e;
          (void)hipStreamSynchronize(st.stream);
          retire_step(sv, st, false);
          sv->cv_done.notify_all();
        }
        progressed = true;
      }
    }
    if (sv->stop) {
      bool any_busy = false;
      for (Step& st : sv->steps) any_busy = any_busy || st.busy;
      if (!any_busy && sv->queue.empty()) break;
    }
    if (!progressed) {
      bool any_busy = false;
      for (Step& st : sv->steps) any_busy = any_busy || st.busy;
      if (any_busy) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      } else {
        sv->cv_work.wait_for(lk, std::chrono::milliseconds(5));
      }
    }
  }
}

void destroy_server(hrt_server* sv) {
  for (Step& st : sv->steps) {
    if (st.done) (void)hipEventDestroy(st.done);
    if (st.stream) (void)hipStreamDestroy(st.stream);
    if (st.tab_host) (void)hipHostFree(st.tab_host);
    if (st.tab_dev) (void)hipFree(st.tab_dev);
    if (st.spill) (void)hipFree(st.spill);
  }
  if (sv->arena_dev) (void)hipFree(sv->arena_dev);
  for (ServedScene& sc : sv->scenes) hrt::scene_serve_ref(sc.scene, -1);
  delete sv;
}

// Where a request buffer lives: 0 pinned host memory, 1 device memory of the
// server's device, -1 neither (pageable host memory, another device).
int buffer_kind(const void* p, int device) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  if (at.type == hipMemoryTypeHost) return 0;
  if (at.type == hipMemoryTypeDevice && at.device == device) return 1;
  return -1;
}

}  // namespace

extern "C" {

hrt_status hrt_host_alloc(size_t bytes, void** out) {
  if (!out) return fail(HRT_ERR_INVALID_ARG, "out is null");
  *out = nullptr;
  hipError_t e = hipHostMalloc(out, std::max<size_t>(bytes, 16), hipHostMallocDefault);
  if (e != hipSuccess) return fail_hip(e, "pinned host allocation");
  return HRT_OK;
}

void hrt_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

hrt_status hrt_server_create(hrt_ctx* ctx, uint64_t budgetBytes, uint32_t stepsInFlight, hrt_server** out) {
  using namespace hrt;
  if (!ctx || !out) return fail(HRT_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (budgetBytes < 4096) return fail(HRT_ERR_INVALID_ARG, "budget below 4 KiB");
  if (stepsInFlight == 0 || stepsInFlight > 8) return fail(HRT_ERR_INVALID_ARG, "1 to 8 steps in flight");
  hrt_server* sv = new (std::nothrow) hrt_server();
  if (!sv) return fail(HRT_ERR_OUT_OF_MEMORY, "host allocation failed");
  sv->ctx = ctx;
  sv->budget = align_up(budgetBytes, 256);
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipMalloc(&sv->arena_dev, sv->budget);
  // steps in flight share the chip: each gets its share of the resident waves
  const uint32_t resident[kKinds] = {resident_grid(reinterpret_cast<const void*>(&k_serve), kWaveS),
                                     serve_any4_resident_grid(true), serve_any4_resident_grid(false)};
  uint32_t gmax = 1;
  for (int t = 0; t < kKinds; t++) {
    sv->grid[t] = std::max<uint32_t>(1, resident[t] / stepsInFlight);
    gmax = std::max(gmax, sv->grid[t]);
  }
  // spill words per lane: the deeper of the BVH2 and BVH4 stacks beyond LDS
  sv->spill_words = size_t(gmax) * kWaveS * size_t(std::max(kStackDepth, kStackDepth4) - kShortStack);
  sv->steps.resize(stepsInFlight);
  for (Step& st : sv->steps) {
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&st.done, hipEventDisableTiming);
    if (e == hipSuccess)
      e = hipHostMalloc(reinterpret_cast<void**>(&st.tab_host), kKinds * kTabBytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&st.tab_dev, kKinds * kTabBytes);
    if (e == hipSuccess) e = hipMalloc(&st.spill, 4 * sv->spill_words);
  }
  if (e != hipSuccess) {
    destroy_server(sv);
    return fail_hip(e, "server allocation");
  }
  sv->arena.init(sv->budget);
  sv->worker = std::thread(worker_loop, sv);
  *out = sv;
  return HRT_OK;
}

void hrt_server_destroy(hrt_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  (void)hipSetDevice(sv->ctx->device);
  destroy_server(sv);
}

hrt_status hrt_server_add_scene(hrt_server* sv, hrt_scene* scene, uint32_t* sceneId) {
  if (!sv || !scene) return fail(HRT_ERR_INVALID_ARG, "null argument");
  const hrt::Blas* b = nullptr;
  hrt_status st = hrt::scene_single_blas(scene, &b);
  if (st != HRT_OK) return st;
  if (hrt::scene_triangle_test(scene) != HRT_TRI_MT)
    return fail(HRT_ERR_UNSUPPORTED, "the server traces with the reference's triangle test");
  ServedScene sc;
  sc.scene = scene;
  sc.pairs = b->pairs;
  sc.tris = b->tris;
  sc.any = hrt::scene_any_hit_tree(scene);
  std::lock_guard<std::mutex> lk(sv->mu);
  hrt::scene_serve_ref(scene, +1);
  sv->scenes.push_back(sc);
  if (sceneId) *sceneId = uint32_t(sv->scenes.size() - 1);
  return HRT_OK;
}

hrt_status hrt_server_submit(hrt_server* sv, uint32_t sceneId, int anyHit, const float* rays, size_t n, void* out,
                             uint64_t* requestId) {
  if (!sv || !requestId) return fail(HRT_ERR_INVALID_ARG, "null argument");
  if (n == 0 || n > 0x7FFFFFFFull) return fail(HRT_ERR_INVALID_ARG, "1 to 2^31-1 rays per request");
  if (!rays || !out) return fail(HRT_ERR_INVALID_ARG, "rays/out is null");
  const int kr = buffer_kind(rays, sv->ctx->device), ko = buffer_kind(out, sv->ctx->device);
  if (kr < 0 || ko < 0)
    return fail(HRT_ERR_INVALID_ARG,
                "request buffers must be pinned host memory (hrt_host_alloc) or device memory of the server's device");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (sceneId >= sv->scenes.size()) return fail(HRT_ERR_INVALID_ARG, "no such scene");
  if (sv->fault != hipSuccess) return fail_hip(sv->fault, "server");
  Request r;
  r.id = sv->next_id++;
  r.scene = sceneId;
  r.any = anyHit ? 1 : 0;
  r.rays = rays;
  r.out = out;
  r.rays_dev = kr == 1;
  r.out_dev = ko == 1;
  r.n = uint32_t(n);
  r.bytes = staging_bytes(r);
  if (r.bytes > sv->budget) return fail(HRT_ERR_OUT_OF_MEMORY, "request larger than the server's HBM budget");
  r.t_submit = hrt::Clock::now();
  sv->reqs[r.id] = r;
  sv->queue.push_back(r.id);
  *requestId = r.id;
  sv->cv_work.notify_one();
  return HRT_OK;
}

// 1 = done (results in the caller's buffer), 0 = pending; negative = failed.
int hrt_server_poll(hrt_server* sv, uint64_t requestId) {
  if (!sv) return -1;
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(requestId);
  if (it == sv->reqs.end()) return -1;
  return it->second.state == 2 ? 1 : (it->second.state == 3 ? -2 : 0);
}

hrt_status hrt_server_wait(hrt_server* sv, uint64_t requestId, double* latencyMs, double* queueMs) {
  if (!sv) return fail(HRT_ERR_INVALID_ARG, "server is null");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(requestId);
  if (it == sv->reqs.end()) return fail(HRT_ERR_INVALID_ARG, "no such request");
  sv->cv_done.wait(lk, [&] { return it->second.state >= 2; });
  const Request r = it->second;
  sv->reqs.erase(it);
  if (r.state == 3) return fail_hip(sv->fault == hipSuccess ? hipErrorUnknown : sv->fault, "request");
  using ms = std::chrono::duration<double, std::milli>;
  if (latencyMs) *latencyMs = ms(r.t_done - r.t_submit).count();
  if (queueMs) *queueMs = ms(r.t_admit - r.t_submit).count();
  return HRT_OK;
}

// out[0..5]: completed requests, rays, steps, mean requests per step x 1000,
// peak bytes in flight, budget bytes.
hrt_status hrt_server_tree_rays(hrt_server* sv, uint64_t out[3]) {
  if (!sv || !out) return fail(HRT_ERR_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  for (int t = 0; t < kKinds; t++) out[t] = sv->kind_rays[t];
  return HRT_OK;
}

hrt_status hrt_server_stats(hrt_server* sv, uint64_t out[6]) {
  if (!sv || !out) return fail(HRT_ERR_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  out[0] = sv->completed;
  out[1] = sv->rays_done;
  out[2] = sv->nsteps;
  out[3] = sv->nsteps ? (1000 * sv->step_reqs) / sv->nsteps : 0;
  out[4] = sv->peak_in_flight;
  out[5] = sv->budget;
  return HRT_OK;
}

}  // extern "C"
