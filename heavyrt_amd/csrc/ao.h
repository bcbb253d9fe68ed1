// ao.h -- ambient-occlusion ray generation, shared by the kernel that writes
// AO rays to HBM (rtao.hip) and the traversal that makes them in registers
// (trace4.hip).  Both call the same functions, so a ray made in either place
// is the same bits.
//
// The reference renderer builds AO rays one at a time on the host
// (RTAO::kernel_InitAORay, examples/02_rtao/rtao.cpp:147-165, with
// MapSampleToCosineDistribution / OffsRayPos, examples/render_common.h:56-112):
// cosine-weighted directions about the surface normal, the origin pushed off
// the surface by epsilonOfPos(hit) along the normal, tNear = 0, tFar = the AO
// radius.  Here a pixel's frame (offset origin and viewer-facing geometric
// normal of its primary hit; hit point = origin + 0.9999995 * t * dir,
// rtao.cpp:92) is made once, and sample k of the whole image draws two PCG32
// numbers with sequence k, so the rays do not depend on the launch shape.
// Pixels whose primary ray missed get an empty ray (tNear = tFar = 0) that
// can hit nothing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hrt_device.h"

namespace hrt {

constexpr float kAoGEps = 2e-5f;  // render_common.h:88

// One pixel's AO frame: org.xyz = hit point pushed off the surface, org.w = 1
// for a hit and 0 for a miss; n.xyz = unit geometric normal facing the viewer,
// n.w = -1 / (sign(n.z) + n.z), the factor of the orthonormal basis about n
// (made once per pixel instead of once per sample).
struct __attribute__((aligned(32))) AoFrame {
  float4 org;
  float4 n;
};

__device__ inline uint32_t ao_pcg32_step(uint64_t& state, uint64_t inc) {
  const uint64_t old = state;
  state = old * 6364136223846793005ull + inc;
  const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = uint32_t(old >> 59u);
  return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

__device__ inline AoFrame ao_frame(const Ray* __restrict__ prim_rays, const uint2* __restrict__ prim_hits,
                                   const float4* __restrict__ verts, const uint32_t* __restrict__ idx,
                                   uint32_t ntris, uint64_t pix) {
  AoFrame f;
  const uint2 h = prim_hits[pix];
  if (h.y >= ntris) {  // miss (or an id outside this geometry)
    f.org = make_float4(0.f, 0.f, 0.f, 0.f);
    f.n = make_float4(0.f, 0.f, 0.f, 0.f);
    return f;
  }
  const Ray p = prim_rays[pix];
  const float t = __uint_as_float(h.x);
  const float3 o = make_float3(p.pos_near.x, p.pos_near.y, p.pos_near.z);
  const float3 d = make_float3(p.dir_far.x, p.dir_far.y, p.dir_far.z);
  const float tt = 0.9999995f * t;
  const float3 hp = make_float3(o.x + tt * d.x, o.y + tt * d.y, o.z + tt * d.z);
  const float4 A = verts[idx[3 * h.y + 0]], B = verts[idx[3 * h.y + 1]], C = verts[idx[3 * h.y + 2]];
  const float3 e1 = make_float3(B.x - A.x, B.y - A.y, B.z - A.z);
  const float3 e2 = make_float3(C.x - A.x, C.y - A.y, C.z - A.z);
  float3 n = make_float3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
  const float inv_len = 1.0f / sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
  n = make_float3(n.x * inv_len, n.y * inv_len, n.z * inv_len);
  if (n.x * d.x + n.y * d.y + n.z * d.z > 0.0f) n = make_float3(-n.x, -n.y, -n.z);
  const float basis_a = -1.0f / (copysignf(1.0f, n.z) + n.z);
  // OffsRayPos: the directions lie in n's hemisphere, so the offset is along +n
  const float m = fmaxf(fmaxf(fabsf(hp.x), fmaxf(fabsf(hp.y), fabsf(hp.z))), 2.0f * kAoGEps) * kAoGEps;
  f.org = make_float4(hp.x + m * n.x, hp.y + m * n.y, hp.z + m * n.z, 1.0f);
  f.n = make_float4(n.x, n.y, n.z, basis_a);
  return f;
}

// sin(2 pi x) and cos(2 pi x) for x in [0, 1).  The nearest quarter turn
// q = rint(4x) is split off exactly (x is a multiple of 2^-24), and the rest
// f = 4x - q in [-1/2, 1/2] becomes theta = f pi/2 in [-pi/4, pi/4] (pi/2 as
// a two-float sum); the Taylor series of sin and cos cut after theta^9 and
// theta^10 (remainders below 2e-9) are evaluated by Horner's rule, and the
// quarter turn picks and signs the two.  Checked against double precision
// on all 2^24 inputs: at most 1.5 ulp (7e-8 absolute).
__device__ inline void ao_sincos_turns(float x, float& s, float& c) {
  const float q = rintf(4.0f * x);
  const float f = 4.0f * x - q;
  constexpr float kHalfPiHi = 1.57079637050628662109375f;          // float(pi / 2)
  constexpr float kHalfPiLo = -4.371138828673792886547744274139404296875e-8f;  // float(pi / 2 - kHalfPiHi)
  const float t = __builtin_fmaf(f, kHalfPiHi, f * kHalfPiLo);
  const float t2 = t * t;
  float sp = __builtin_fmaf(t2, 1.0f / 362880.0f, -1.0f / 5040.0f);
  sp = __builtin_fmaf(sp, t2, 1.0f / 120.0f);
  sp = __builtin_fmaf(sp, t2, -1.0f / 6.0f);
  const float s0 = __builtin_fmaf(t * t2, sp, t);
  float cp = __builtin_fmaf(t2, -1.0f / 3628800.0f, 1.0f / 40320.0f);
  cp = __builtin_fmaf(cp, t2, -1.0f / 720.0f);
  cp = __builtin_fmaf(cp, t2, 1.0f / 24.0f);
  cp = __builtin_fmaf(cp, t2, -0.5f);
  const float c0 = __builtin_fmaf(cp, t2, 1.0f);
  const int qi = int(q) & 3;
  s = (qi & 1) ? c0 : s0;
  c = (qi & 1) ? s0 : c0;
  if (qi == 2 || qi == 3) s = -s;
  if (qi == 1 || qi == 2) c = -c;
}

// Sample k of the image (k = pixel * spp + s) from its pixel's frame.
__device__ inline Ray ao_ray(const AoFrame& f, uint64_t k, float tmax, uint64_t seed) {
  Ray r;
  if (f.org.w == 0.0f) {
    r.pos_near = make_float4(0.f, 0.f, 0.f, 0.f);
    r.dir_far = make_float4(0.f, 1.f, 0.f, 0.f);
    return r;
  }
  const float3 n = make_float3(f.n.x, f.n.y, f.n.z);
  // PCG32(seed, sequence k): pcg32_srandom_r then two draws
  const uint64_t inc = (k << 1u) | 1u;
  uint64_t st = 0;
  (void)ao_pcg32_step(st, inc);
  st += seed;
  (void)ao_pcg32_step(st, inc);
  const float r1 = float(ao_pcg32_step(st, inc) >> 8) * (1.0f / 16777216.0f);
  const float r2 = float(ao_pcg32_step(st, inc) >> 8) * (1.0f / 16777216.0f);
  // cosine-weighted hemisphere (power 1 lobe, render_common.h:62-70)
  float sphi, cphi;
  ao_sincos_turns(r1, sphi, cphi);
  const float cth = sqrtf(1.0f - r2);
  const float sth = sqrtf(r2);
  // orthonormal basis about n (Duff et al., "Building an Orthonormal Basis, Revisited", JCGT 2017)
  const float sgn = copysignf(1.0f, n.z);
  const float a = f.n.w;
  const float b = n.x * n.y * a;
  const float3 tx = make_float3(1.0f + sgn * n.x * n.x * a, sgn * b, -sgn * n.x);
  const float3 ty = make_float3(b, sgn + n.y * n.y * a, -n.y);
  const float lx = sth * cphi, ly = sth * sphi;
  r.pos_near = make_float4(f.org.x, f.org.y, f.org.z, 0.0f);
  r.dir_far = make_float4(tx.x * lx + ty.x * ly + n.x * cth, tx.y * lx + ty.y * ly + n.y * cth,
                          tx.z * lx + ty.z * ly + n.z * cth, tmax);
  return r;
}

// AO samples are traced in direction order inside runs of kAoOrderRun
// consecutive samples: the wave that takes a run's next rays then holds rays
// of one direction bin from neighbouring pixels, so their node and triangle
// fetches share cache lines and their traversals take similar paths.
// Any-hit answers do not depend on which lane traces which ray.
constexpr uint32_t kAoOrderRun = 1024;
constexpr uint32_t kAoBins = 24;
// 24 direction bins: the cube face of the largest |d| component (the lower
// axis on ties; 6 faces) times the signs of the two other components.
__device__ inline uint32_t ao_dir_bin(float3 d) {
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  uint32_t face;
  float p, q;
  if (ax >= ay && ax >= az) {
    face = d.x < 0.0f ? 1u : 0u;
    p = d.y, q = d.z;
  } else if (ay >= az) {
    face = d.y < 0.0f ? 3u : 2u;
    p = d.x, q = d.z;
  } else {
    face = d.z < 0.0f ? 5u : 4u;
    p = d.x, q = d.y;
  }
  return face * 4u + (p >= 0.0f ? 2u : 0u) + (q >= 0.0f ? 1u : 0u);
}

}  // namespace hrt
