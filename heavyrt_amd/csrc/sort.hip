// sort.hip -- stable LSD radix sort of (key, value) u32 pairs, onesweep
// style (one histogram pass for every digit, then one scatter launch per
// digit with decoupled look-back between tiles; Adinets & Merrill 2022).
//
// The LBVH order is (Morton code, triangle id) (lbvh.cpp:755 sorts on the
// code only and leaves equal codes in an unspecified order; fixing the order
// to ascending ids makes the build deterministic).  The ids start ascending,
// so a stable sort on the code yields exactly that order.
//
// Per digit pass, one workgroup per 4,096-key tile:
//   1. takes the next tile number from a counter (tiles start in order, so
//      every tile it waits for below is already running);
//   2. ranks its keys per wave with eight ballots per 64 keys (stable: equal
//      digits keep index order) and counts each digit;
//   3. publishes its 256 digit counts, then walks back over earlier tiles'
//      published counts until it meets an inclusive prefix, and publishes its
//      own inclusive prefix (count and flag share one word, so relaxed
//      device-scope atomics suffice: no cache flushes);
//   4. sorts the tile by digit in LDS, then writes each digit's keys to
//      consecutive global positions, so stores coalesce along digit runs.
#include <hip/hip_runtime.h>

#include "hrt_internal.h"

namespace hrt {
namespace {

constexpr int kOsThreads = 256;
constexpr int kOsWaves = kOsThreads / 64;
constexpr int kOsRounds = 16;
constexpr uint32_t kOsTile = kOsWaves * kOsRounds * 64;  // 4096 keys
constexpr int kOsDigits = 256;
constexpr int kOsMaxPasses = 4;
constexpr uint32_t kOsHistTile = 4 * kOsTile;  // keys per workgroup in the histogram pass

// status word: flag in bits 31:30, count in bits 29:0
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1u;
// polls of one predecessor before the pass gives up and flags an error
// (a tile publishes its counts without waiting for anything, so this is
// only a guard that lets every wave finish)
constexpr uint32_t kSpinLimit = 1u << 22;

// scratch words: per-pass digit offsets, per-pass tile counters, error flag,
// then per-pass tile status
constexpr size_t kOffWords = size_t(kOsMaxPasses) * kOsDigits;
constexpr size_t kCtrWords = 64;
constexpr size_t kHeadWords = kOffWords + kCtrWords;

inline uint32_t tiles_for(uint64_t n) { return uint32_t((n + kOsTile - 1) / kOsTile); }

// Exclusive scan over the 256 threads' values (one per thread).
__device__ inline uint32_t block_scan_256(uint32_t x, uint32_t* tmp) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t incl = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  if (lane == 63) tmp[w] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (int ww = 0; ww < w; ww++) before += tmp[ww];
  __syncthreads();
  return before + incl - x;
}

// All digit histograms in one read of the keys.
__global__ __launch_bounds__(kOsThreads) void k_os_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift0,
                                                        int npasses, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kOsWaves][kOsMaxPasses][kOsDigits];
  for (int k = threadIdx.x; k < kOsWaves * kOsMaxPasses * kOsDigits; k += kOsThreads) (&h[0][0][0])[k] = 0u;
  __syncthreads();
  const int w = threadIdx.x >> 6;
  const uint64_t base = uint64_t(blockIdx.x) * kOsHistTile;
  const uint64_t end = base + kOsHistTile < n ? base + kOsHistTile : n;
  for (uint64_t i = base + threadIdx.x; i < end; i += kOsThreads) {
    const uint32_t k = keys[i];
    for (int p = 0; p < npasses; p++) atomicAdd(&h[w][p][(k >> (shift0 + 8 * p)) & 0xFFu], 1u);
  }
  __syncthreads();
  for (int p = 0; p < npasses; p++) {
    uint32_t v = 0;
#pragma unroll
    for (int ww = 0; ww < kOsWaves; ww++) v += h[ww][p][threadIdx.x];
    if (v) atomicAdd(&hist[p * kOsDigits + threadIdx.x], v);
  }
}

// Histograms -> exclusive digit offsets, one workgroup per pass.
__global__ __launch_bounds__(kOsThreads) void k_os_offsets(uint32_t* __restrict__ hist) {
  __shared__ uint32_t tmp[kOsWaves];
  uint32_t* h = hist + blockIdx.x * kOsDigits;
  const uint32_t x = h[threadIdx.x];
  const uint32_t ex = block_scan_256(x, tmp);
  h[threadIdx.x] = ex;
}

__global__ __launch_bounds__(kOsThreads) void k_os_pass(const uint32_t* __restrict__ kin,
                                                        const uint32_t* __restrict__ vin,
                                                        uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                        uint32_t n, int shift, const uint32_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ counter, uint32_t* status,
                                                        uint32_t* __restrict__ err) {
  __shared__ uint32_t s_tile;
  __shared__ uint32_t wcnt[kOsWaves][kOsDigits];
  __shared__ uint32_t tstart[kOsDigits];
  __shared__ uint32_t gbase[kOsDigits];
  __shared__ uint32_t scan_tmp[kOsWaves];
  __shared__ uint32_t stage_k[kOsTile];
  __shared__ uint32_t stage_v[kOsTile];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
  for (int k = threadIdx.x; k < kOsWaves * kOsDigits; k += kOsThreads) (&wcnt[0][0])[k] = 0u;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t tile_base = uint64_t(tile) * kOsTile;

  // 2. load and rank; wave w owns keys [w * 1024, w * 1024 + 1024) of the tile
  const uint64_t base = tile_base + uint64_t(w) * (kOsRounds * 64);
  uint32_t key[kOsRounds], val[kOsRounds], pos[kOsRounds];
#pragma unroll
  for (int r = 0; r < kOsRounds; r++) {
    const uint64_t i = base + uint64_t(r) * 64 + lane;
    const bool valid = i < n;
    key[r] = valid ? kin[i] : 0u;
    val[r] = valid ? vin[i] : 0u;
  }
#pragma unroll
  for (int r = 0; r < kOsRounds; r++) {
    const uint64_t i = base + uint64_t(r) * 64 + lane;
    const bool valid = i < n;
    const uint32_t digit = (key[r] >> shift) & 0xFFu;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (digit >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = lanes_below(peers);
    const uint32_t cnt = __popcll(peers);
    const uint32_t before = wcnt[w][digit];
    __builtin_amdgcn_wave_barrier();
    if (valid && rank == 0) wcnt[w][digit] = before + cnt;
    __builtin_amdgcn_wave_barrier();
    pos[r] = before + rank;
  }
  __syncthreads();

  // 3. per-wave offsets inside the tile, tile totals, look-back
  const int d = threadIdx.x;  // kOsThreads == kOsDigits
  uint32_t tot = 0;
#pragma unroll
  for (int ww = 0; ww < kOsWaves; ww++) {
    const uint32_t c = wcnt[ww][d];
    wcnt[ww][d] = tot;
    tot += c;
  }
  uint32_t* my = status + uint64_t(tile) * kOsDigits + d;
  uint32_t excl = 0;
  if (tile == 0) {
    __hip_atomic_store(my, kFlagPrefix | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_store(my, kFlagAgg | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t j = tile - 1, spins = 0;
    while (true) {
      const uint32_t v =
          __hip_atomic_load(status + uint64_t(j) * kOsDigits + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t flag = v & ~kCountMask;
      if (flag == 0u) {
        if (++spins > kSpinLimit) {
          __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      excl += v & kCountMask;
      if (flag == kFlagPrefix || j == 0) break;
      j--;
      spins = 0;
    }
    __hip_atomic_store(my, kFlagPrefix | (excl + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  gbase[d] = offsets[d] + excl;
  tstart[d] = block_scan_256(tot, scan_tmp);
  __syncthreads();

  // 4. digit order in LDS, then runs of equal digits to consecutive addresses
#pragma unroll
  for (int r = 0; r < kOsRounds; r++) {
    const uint64_t i = base + uint64_t(r) * 64 + lane;
    if (i < n) {
      const uint32_t digit = (key[r] >> shift) & 0xFFu;
      const uint32_t lp = tstart[digit] + wcnt[w][digit] + pos[r];
      stage_k[lp] = key[r];
      stage_v[lp] = val[r];
    }
  }
  __syncthreads();
  const uint32_t count = uint32_t(n - tile_base < kOsTile ? n - tile_base : kOsTile);
  for (uint32_t k = threadIdx.x; k < count; k += kOsThreads) {
    const uint32_t kk = stage_k[k];
    const uint32_t digit = (kk >> shift) & 0xFFu;
    const uint32_t g = gbase[digit] + (k - tstart[digit]);
    kout[g] = kk;
    vout[g] = stage_v[k];
  }
}

__global__ __launch_bounds__(256) void k_os_fill(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256) p[i] = 0u;
}

}  // namespace

size_t radix_sort_scratch_words(uint64_t n, int npasses) {
  return kHeadWords + size_t(npasses) * tiles_for(n) * kOsDigits;
}

hipError_t radix_sort(uint32_t* ka, uint32_t* va, uint32_t* kb, uint32_t* vb, uint64_t n, int shift0, int npasses,
                      uint32_t* scratch, uint32_t* err, hipStream_t s) {
  if (n == 0 || npasses == 0) return hipSuccess;
  if (npasses < 0 || npasses > kOsMaxPasses || shift0 < 0 || shift0 + 8 * npasses > 32 || n > kCountMask)
    return hipErrorInvalidValue;
  const uint32_t ntiles = tiles_for(n);
  const size_t words = radix_sort_scratch_words(n, npasses);
  const uint32_t fill_grid = uint32_t(std::min<uint64_t>(1024, (words + 255) / 256));
  k_os_fill<<<fill_grid, 256, 0, s>>>(scratch, words);
  uint32_t* offsets = scratch;
  uint32_t* counters = scratch + kOffWords;
  uint32_t* status = scratch + kHeadWords;
  k_os_hist<<<uint32_t((n + kOsHistTile - 1) / kOsHistTile), kOsThreads, 0, s>>>(ka, uint32_t(n), shift0, npasses,
                                                                                  offsets);
  k_os_offsets<<<npasses, kOsThreads, 0, s>>>(offsets);
  uint32_t *kin = ka, *vin = va, *kout = kb, *vout = vb;
  for (int p = 0; p < npasses; p++) {
    k_os_pass<<<ntiles, kOsThreads, 0, s>>>(kin, vin, kout, vout, uint32_t(n), shift0 + 8 * p,
                                            offsets + p * kOsDigits, counters + p,
                                            status + size_t(p) * ntiles * kOsDigits, err);
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  return hipGetLastError();
}

}  // namespace hrt
