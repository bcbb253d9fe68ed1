// eye.hip -- the 01_eye_rays renderer's ray generation and colouring on the
// device.
//
// The reference renders one pixel per call (EyeRayCaster::CastRaySingle,
// examples/01_eye_rays/eye_ray.cpp:13-24): thread tid reads its pixel from
// m_packedXY, which kernel_PackXY (eye_ray.cpp:133-138) fills in
// SuperBlockIndex2DOpt order (eye_ray.cpp:114-131: 4x4 pixel blocks, 2x2
// blocks per 8x8 superblock, superblocks row by row), builds its ray with
// EyeRayDirNormalized + transform_ray3f (eye_ray.cpp:26-47,
// render_common.h:6-36) and colours the pixel from the hit's primId
// (eye_ray.cpp:50-75).  Here one lane makes one ray directly in HBM, in that
// same order for hrt_eye_make_rays (64 consecutive rays are one 8x8
// superblock).  The whole frame on the device (hrt_eye_render) only owes the
// image, so it makes, traces and colours its rays in the order that traces
// fastest: XCD stripes (xcd_stripe_pixel).  On the 512x512 eye frame the
// trace took 0.126 ms that way, 0.145 in row order, 0.164 in superblocks.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "hrt_internal.h"

namespace hrt {
namespace {

constexpr int kBlock = 256;

// Thread tid -> pixel: the inverse of SuperBlockIndex2DOpt for a width that
// is a multiple of 8.
__device__ inline uint2 superblock_pixel(uint32_t tid, uint32_t width) {
  const uint32_t wblocks_h = width >> 3;
  const uint32_t sb = tid >> 6, in_h = (tid >> 4) & 3u, in_b = tid & 15u;
  const uint32_t bx = (sb % wblocks_h) * 2u + (in_h & 1u);
  const uint32_t by = (sb / wblocks_h) * 2u + (in_h >> 1);
  return make_uint2(bx * 4u + (in_b & 3u), by * 4u + (in_b >> 2));
}

// Thread tid -> pixel in XCD stripes: the nearest-hit kernel runs rays
// [256 b, 256 b + 256) as workgroup b, which the dispatcher hands to XCD
// b mod 8, so ray tid is given a pixel of the rows r with (r / 4) mod 8 equal
// to that XCD (4-row stripes, row order within each XCD's share).  Each XCD's
// L2 then serves one eighth of the screen while the stripes keep every XCD's
// share spread over the whole frame (contiguous eighths leave some XCDs the
// costly regions).  For a height that is a multiple of 32 and a width that
// is a multiple of 64 (each XCD's share is then a whole number of
// workgroups); rows otherwise.  A permutation of the pixels either way.
constexpr uint32_t kTraceGroup = 256;  // lanes per workgroup of the nearest-hit kernel
constexpr uint32_t kStripeRows = 4;
__device__ inline uint2 xcd_stripe_pixel(uint32_t tid, uint32_t width, uint32_t height) {
  if (height % (kStripeRows * kXcds) != 0 || width % (kTraceGroup / kStripeRows) != 0)
    return make_uint2(tid % width, tid / width);
  const uint32_t b = tid / kTraceGroup, x = b % kXcds;
  const uint32_t k = (b / kXcds) * kTraceGroup + tid % kTraceGroup;  // place in XCD x's share
  const uint32_t rk = k / width;
  const uint32_t row = (rk / kStripeRows) * (kStripeRows * kXcds) + x * kStripeRows + rk % kStripeRows;
  return make_uint2(k % width, row);
}

// Thread tid -> pixel in superblock order, or in XCD stripes (kStripes).
template <bool kStripes>
__device__ inline uint2 thread_pixel(uint32_t tid, uint32_t width, uint32_t height) {
  return kStripes ? xcd_stripe_pixel(tid, width, height) : superblock_pixel(tid, width);
}

// LiteMath float4x4 * float4 with a column-major matrix: each row's four
// products summed left to right.
__device__ inline float4 mat_vec(const EyeCamera::Mat& m, float4 v) {
  const float* a = m.m;
  return make_float4(a[0] * v.x + a[4] * v.y + a[8] * v.z + a[12] * v.w,
                     a[1] * v.x + a[5] * v.y + a[9] * v.z + a[13] * v.w,
                     a[2] * v.x + a[6] * v.y + a[10] * v.z + a[14] * v.w,
                     a[3] * v.x + a[7] * v.y + a[11] * v.z + a[15] * v.w);
}

// normalize(u) = u / length(u), length = sqrt(dot(u, u))
__device__ inline float3 normalize3(float3 u) {
  const float len = sqrtf(u.x * u.x + u.y * u.y + u.z * u.z);
  return make_float3(u.x / len, u.y / len, u.z / len);
}

template <bool kStripes>
__global__ __launch_bounds__(kBlock) void k_eye_rays(EyeCamera cam, uint32_t width, uint32_t height,
                                                     Ray* __restrict__ out, uint32_t* __restrict__ packed_xy) {
  const uint32_t tid = blockIdx.x * kBlock + threadIdx.x;
  if (tid >= width * height) return;
  const uint2 px = thread_pixel<kStripes>(tid, width, height);
  const float x = (float(px.x) + 0.5f) / float(width);
  const float y = (float(px.y) + 0.5f) / float(height);
  // EyeRayDirNormalized
  float4 pos = make_float4(2.0f * x - 1.0f, -2.0f * y + 1.0f, 0.0f, 1.0f);
  pos = mat_vec(cam.proj_inv, pos);
  const float w = pos.w;
  pos = make_float4(pos.x / w, pos.y / w, pos.z / w, pos.w / w);
  float3 dir = normalize3(make_float3(pos.x, pos.y, pos.z));
  // transform_ray3f from the camera origin
  const float3 o = make_float3(0.0f, 0.0f, 0.0f);
  const float4 p1 = mat_vec(cam.world_view_inv, make_float4(o.x, o.y, o.z, 1.0f));
  const float3 o2 = make_float3(o.x + 100.0f * dir.x, o.y + 100.0f * dir.y, o.z + 100.0f * dir.z);
  const float4 p2 = mat_vec(cam.world_view_inv, make_float4(o2.x, o2.y, o2.z, 1.0f));
  dir = normalize3(make_float3(p2.x - p1.x, p2.y - p1.y, p2.z - p1.z));
  Ray r;
  r.pos_near = make_float4(p1.x, p1.y, p1.z, 0.0f);
  r.dir_far = make_float4(dir.x, dir.y, dir.z, FLT_MAX);
  out[tid] = r;
  if (packed_xy) packed_xy[tid] = (px.y << 16) | px.x;
}

// out_color[y * width + x] = 0 on a miss, else palette[primId % size]
template <bool kStripes>
__global__ __launch_bounds__(kBlock) void k_eye_colors(const uint2* __restrict__ hits, uint32_t width,
                                                       uint32_t height, EyePalette pal,
                                                       uint32_t* __restrict__ out_color) {
  const uint32_t tid = blockIdx.x * kBlock + threadIdx.x;
  if (tid >= width * height) return;
  const uint2 px = thread_pixel<kStripes>(tid, width, height);
  const uint32_t prim = hits[tid].y;
  out_color[px.y * width + px.x] = prim == 0xFFFFFFFFu ? 0u : pal.c[prim % pal.n];
}

// The 02_rtao frame's colour (RTAO::kernel_AO2Color and floatToGrayRGBA,
// examples/02_rtao/rtao.cpp:15-24, 62-72): 0 where the eye ray missed, else
// the occluded fraction of the pixel's spp AO rays, raised to `power`,
// clamped to [0, 1] and truncated to a gray level with alpha 255.
__global__ __launch_bounds__(kBlock) void k_ao_colors(const uint2* __restrict__ hits,
                                                      const uint8_t* __restrict__ occluded, uint32_t width,
                                                      uint32_t height, uint32_t spp, float power,
                                                      uint32_t* __restrict__ out_color) {
  const uint32_t tid = blockIdx.x * kBlock + threadIdx.x;
  if (tid >= width * height) return;
  const uint2 px = superblock_pixel(tid, width);
  uint32_t c = 0u;
  if (hits[tid].y != 0xFFFFFFFFu) {
    uint32_t count = 0;
    for (uint32_t k = 0; k < spp; k++) count += occluded[uint64_t(tid) * spp + k] ? 1u : 0u;
    float v = float(count) * (1.0f / float(spp));
    if (power != 1.0f) v = powf(v, power);
    const uint32_t g = uint32_t(fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f);
    c = g | (g << 8) | (g << 16) | (255u << 24);
  }
  out_color[px.y * width + px.x] = c;
}

}  // namespace

hipError_t ao_colors(const void* hits_compact, const uint8_t* occluded, uint32_t width, uint32_t height,
                     uint32_t spp, float power, uint32_t* out_color, hipStream_t stream) {
  const uint64_t n = uint64_t(width) * height;
  if (n == 0) return hipSuccess;
  k_ao_colors<<<uint32_t((n + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
      static_cast<const uint2*>(hits_compact), occluded, width, height, spp, power, out_color);
  return hipGetLastError();
}

hipError_t eye_make_rays(const EyeCamera& cam, uint32_t width, uint32_t height, Ray* out, uint32_t* packed_xy,
                         hipStream_t stream, bool stripes) {
  const uint64_t n = uint64_t(width) * height;
  if (n == 0) return hipSuccess;
  const uint32_t blocks = uint32_t((n + kBlock - 1) / kBlock);
  if (stripes) k_eye_rays<true><<<blocks, kBlock, 0, stream>>>(cam, width, height, out, packed_xy);
  else k_eye_rays<false><<<blocks, kBlock, 0, stream>>>(cam, width, height, out, packed_xy);
  return hipGetLastError();
}

hipError_t eye_colors(const void* hits_compact, uint32_t width, uint32_t height, const EyePalette& pal,
                      uint32_t* out_color, hipStream_t stream, bool stripes) {
  const uint64_t n = uint64_t(width) * height;
  if (n == 0) return hipSuccess;
  const uint32_t blocks = uint32_t((n + kBlock - 1) / kBlock);
  const uint2* hits = static_cast<const uint2*>(hits_compact);
  if (stripes) k_eye_colors<true><<<blocks, kBlock, 0, stream>>>(hits, width, height, pal, out_color);
  else k_eye_colors<false><<<blocks, kBlock, 0, stream>>>(hits, width, height, pal, out_color);
  return hipGetLastError();
}

}  // namespace hrt
