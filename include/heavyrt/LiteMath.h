// LiteMath.h -- the small vector/matrix vocabulary that the ray-query
// interface (CrossRT.h) and renderer-style callers use, written for this
// engine.  HeavyRT takes these names from its LiteMath submodule (float3,
// float4, float4x4 column-major, Box4f, dot/cross/normalize, ...); this
// header provides the same names and meanings with plain scalar arithmetic
// evaluated left to right, so code written against the reference interface
// compiles unchanged against the engine.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace LiteMath {

typedef unsigned int uint;

constexpr float MAXFLOAT_ = 3.402823466e+38f;
#ifndef MAXFLOAT
#define MAXFLOAT 3.402823466e+38f
#endif
constexpr float DEG_TO_RAD = 3.14159265358979323846f / 180.0f;
constexpr float M_TWOPI_ = 6.28318530717958647692f;
#ifndef M_TWOPI
#define M_TWOPI 6.28318530717958647692
#endif
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
constexpr float INV_PI = 0.31830988618379067154f;
constexpr float INV_TWOPI = 0.15915494309189533577f;

struct float2 {
  float x, y;
  float2() : x(0), y(0) {}
  float2(float a, float b) : x(a), y(b) {}
  explicit float2(float a) : x(a), y(a) {}
  explicit float2(const float* p) : x(p[0]), y(p[1]) {}
  float& operator[](int i) { return (&x)[i]; }
  float operator[](int i) const { return (&x)[i]; }
};

struct float3 {
  float x, y, z;
  float3() : x(0), y(0), z(0) {}
  float3(float a, float b, float c) : x(a), y(b), z(c) {}
  explicit float3(float a) : x(a), y(a), z(a) {}
  explicit float3(const float* p) : x(p[0]), y(p[1]), z(p[2]) {}
  float& operator[](int i) { return (&x)[i]; }
  float operator[](int i) const { return (&x)[i]; }
};

struct float4 {
  float x, y, z, w;
  float4() : x(0), y(0), z(0), w(0) {}
  float4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
  explicit float4(float a) : x(a), y(a), z(a), w(a) {}
  explicit float4(const float* p) : x(p[0]), y(p[1]), z(p[2]), w(p[3]) {}
  float& operator[](int i) { return (&x)[i]; }
  float operator[](int i) const { return (&x)[i]; }
};

struct int2 { int x, y; int2() : x(0), y(0) {} int2(int a, int b) : x(a), y(b) {} };
struct int3 { int x, y, z; int3() : x(0), y(0), z(0) {} int3(int a, int b, int c) : x(a), y(b), z(c) {} };
// int4(float4) truncates toward zero per component, as a C cast does.
struct int4 {
  int x, y, z, w;
  int4() : x(0), y(0), z(0), w(0) {}
  int4(int a, int b, int c, int d) : x(a), y(b), z(c), w(d) {}
  explicit int4(float4 f) : x(int(f.x)), y(int(f.y)), z(int(f.z)), w(int(f.w)) {}
};
struct uint2 { uint x, y; uint2() : x(0), y(0) {} uint2(uint a, uint b) : x(a), y(b) {} };
struct uint3 { uint x, y, z; uint3() : x(0), y(0), z(0) {} uint3(uint a, uint b, uint c) : x(a), y(b), z(c) {} };
struct uint4 { uint x, y, z, w; uint4() : x(0), y(0), z(0), w(0) {} uint4(uint a, uint b, uint c, uint d) : x(a), y(b), z(c), w(d) {} };
typedef unsigned short ushort;
typedef unsigned char uchar;
struct ushort2 { ushort x, y; ushort2() : x(0), y(0) {} ushort2(ushort a, ushort b) : x(a), y(b) {} };
struct ushort4 {
  ushort x, y, z, w;
  ushort4() : x(0), y(0), z(0), w(0) {}
  ushort4(ushort a, ushort b, ushort c, ushort d) : x(a), y(b), z(c), w(d) {}
};
struct uchar4 {
  uchar x, y, z, w;
  uchar4() : x(0), y(0), z(0), w(0) {}
  uchar4(uchar a, uchar b, uchar c, uchar d) : x(a), y(b), z(c), w(d) {}
};

// --- float3 ---------------------------------------------------------------
inline float3 operator+(float3 a, float3 b) { return float3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline float3 operator-(float3 a, float3 b) { return float3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline float3 operator*(float3 a, float3 b) { return float3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline float3 operator/(float3 a, float3 b) { return float3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline float3 operator*(float3 a, float s) { return float3(a.x * s, a.y * s, a.z * s); }
inline float3 operator*(float s, float3 a) { return float3(s * a.x, s * a.y, s * a.z); }
inline float3 operator/(float3 a, float s) { return float3(a.x / s, a.y / s, a.z / s); }
inline float3 operator-(float3 a) { return float3(-a.x, -a.y, -a.z); }
inline float3& operator+=(float3& a, float3 b) { a = a + b; return a; }
inline float3& operator*=(float3& a, float s) { a = a * s; return a; }
inline float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float3 cross(float3 a, float3 b) {
  return float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length(float3 a) { return std::sqrt(dot(a, a)); }
inline float3 normalize(float3 a) { return a * (1.0f / length(a)); }
inline float3 min(float3 a, float3 b) { return float3(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)); }
inline float3 max(float3 a, float3 b) { return float3(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)); }

// --- float4 ---------------------------------------------------------------
inline float4 operator+(float4 a, float4 b) { return float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
inline float4 operator-(float4 a, float4 b) { return float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
inline float4 operator*(float4 a, float4 b) { return float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
inline float4 operator*(float4 a, float s) { return float4(a.x * s, a.y * s, a.z * s, a.w * s); }
inline float4 operator*(float s, float4 a) { return float4(s * a.x, s * a.y, s * a.z, s * a.w); }
inline float4 operator/(float4 a, float s) { return float4(a.x / s, a.y / s, a.z / s, a.w / s); }
inline float4& operator/=(float4& a, float s) { a = a / s; return a; }
inline float dot(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
inline float4 min(float4 a, float4 b) {
  return float4(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z), std::fmin(a.w, b.w));
}
inline float4 max(float4 a, float4 b) {
  return float4(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z), std::fmax(a.w, b.w));
}
inline float4 operator-(float4 a) { return float4(-a.x, -a.y, -a.z, -a.w); }
inline float4 operator+(float4 a, float s) { return float4(a.x + s, a.y + s, a.z + s, a.w + s); }
inline float4 operator-(float4 a, float s) { return float4(a.x - s, a.y - s, a.z - s, a.w - s); }
inline float4 operator/(float4 a, float4 b) { return float4(a.x / b.x, a.y / b.y, a.z / b.z, a.w / b.w); }
inline float4& operator+=(float4& a, float4 b) { a = a + b; return a; }
inline float4& operator-=(float4& a, float4 b) { a = a - b; return a; }
inline float4& operator*=(float4& a, float s) { a = a * s; return a; }
inline float length(float4 a) { return std::sqrt(dot(a, a)); }
inline float4 normalize(float4 a) { return a * (1.0f / length(a)); }
inline float length3f(float4 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline float4 abs(float4 a) { return float4(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z), std::fabs(a.w)); }
inline float3 abs(float3 a) { return float3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
inline float4 clamp(float4 v, float lo, float hi) {
  return float4(v.x < lo ? lo : (v.x > hi ? hi : v.x), v.y < lo ? lo : (v.y > hi ? hi : v.y),
                v.z < lo ? lo : (v.z > hi ? hi : v.z), v.w < lo ? lo : (v.w > hi ? hi : v.w));
}
inline float3 clamp(float3 v, float lo, float hi) {
  return float3(v.x < lo ? lo : (v.x > hi ? hi : v.x), v.y < lo ? lo : (v.y > hi ? hi : v.y),
                v.z < lo ? lo : (v.z > hi ? hi : v.z));
}
inline float dot(float a, float b) { return a * b; }
inline float3 to_float3(float4 a) { return float3(a.x, a.y, a.z); }
inline float4 to_float4(float3 a, float w) { return float4(a.x, a.y, a.z, w); }
inline float as_float(uint u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint as_uint(float f) { uint u; std::memcpy(&u, &f, 4); return u; }
inline float clamp(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline int clamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline uint clamp(uint v, uint lo, uint hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline float lerp(float a, float b, float t) { return a + t * (b - a); }
inline int sign(int v) { return (v > 0) - (v < 0); }
inline float sign(float v) { return (v > 0.0f) ? 1.0f : ((v < 0.0f) ? -1.0f : 0.0f); }

// --- float4x4, column-major: m_col[c] is column c ------------------------
struct float4x4 {
  float4 m_col[4];
  float4x4() { identity(); }
  void identity() {
    m_col[0] = float4(1, 0, 0, 0);
    m_col[1] = float4(0, 1, 0, 0);
    m_col[2] = float4(0, 0, 1, 0);
    m_col[3] = float4(0, 0, 0, 1);
  }
  float4 get_col(int i) const { return m_col[i]; }
  void set_col(int i, float4 c) { m_col[i] = c; }
  float4 get_row(int i) const { return float4(m_col[0][i], m_col[1][i], m_col[2][i], m_col[3][i]); }
  void set_row(int i, float4 r) { m_col[0][i] = r.x; m_col[1][i] = r.y; m_col[2][i] = r.z; m_col[3][i] = r.w; }
  float& operator()(int row, int col) { return m_col[col][row]; }
  float operator()(int row, int col) const { return m_col[col][row]; }
  const float* data() const { return &m_col[0].x; }
};

inline float4 operator*(const float4x4& m, float4 v) {
  return m.m_col[0] * v.x + m.m_col[1] * v.y + m.m_col[2] * v.z + m.m_col[3] * v.w;
}
// A point (w = 1) through the matrix, without a perspective divide.
inline float3 operator*(const float4x4& m, float3 v) { return to_float3(m * to_float4(v, 1.0f)); }
inline float4 mul(const float4x4& m, float4 v) { return m * v; }
inline float3 mul(const float4x4& m, float3 v) { return m * v; }
inline float4x4 operator*(const float4x4& a, const float4x4& b) {
  float4x4 r;
  for (int c = 0; c < 4; c++) r.m_col[c] = a * b.m_col[c];
  return r;
}
inline float4x4 transpose(const float4x4& m) {
  float4x4 r;
  for (int c = 0; c < 4; c++) r.set_row(c, m.get_col(c));
  return r;
}
inline float4x4 translate4x4(float3 t) {
  float4x4 m;
  m.m_col[3] = float4(t.x, t.y, t.z, 1.0f);
  return m;
}
inline float4x4 scale4x4(float3 s) {
  float4x4 m;
  m.m_col[0].x = s.x;
  m.m_col[1].y = s.y;
  m.m_col[2].z = s.z;
  return m;
}
inline float4x4 rotate4x4X(float a) {
  float4x4 m;
  m.m_col[1] = float4(0, std::cos(a), std::sin(a), 0);
  m.m_col[2] = float4(0, -std::sin(a), std::cos(a), 0);
  return m;
}
inline float4x4 rotate4x4Y(float a) {
  float4x4 m;
  m.m_col[0] = float4(std::cos(a), 0, -std::sin(a), 0);
  m.m_col[2] = float4(std::sin(a), 0, std::cos(a), 0);
  return m;
}
inline float4x4 rotate4x4Z(float a) {
  float4x4 m;
  m.m_col[0] = float4(std::cos(a), std::sin(a), 0, 0);
  m.m_col[1] = float4(-std::sin(a), std::cos(a), 0, 0);
  return m;
}
// General inverse by cofactors (double accumulation).
inline float4x4 inverse4x4(const float4x4& m) {
  double a[16], inv[16];
  for (int c = 0; c < 4; c++)
    for (int r = 0; r < 4; r++) a[4 * r + c] = m(r, c);
  inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
  inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
  inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
  inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
  inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
  inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
  inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
  inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
  inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
  inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
  inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
  inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
  inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
  inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
  inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
  inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
  const double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
  float4x4 r;
  for (int row = 0; row < 4; row++)
    for (int col = 0; col < 4; col++) r(row, col) = float(inv[4 * row + col] / det);
  return r;
}
inline float4x4 lookAt(float3 eye, float3 center, float3 up) {
  const float3 f = normalize(center - eye);
  const float3 s = normalize(cross(f, up));
  const float3 u = cross(s, f);
  float4x4 m;
  m.set_row(0, float4(s.x, s.y, s.z, -dot(s, eye)));
  m.set_row(1, float4(u.x, u.y, u.z, -dot(u, eye)));
  m.set_row(2, float4(-f.x, -f.y, -f.z, dot(f, eye)));
  m.set_row(3, float4(0, 0, 0, 1));
  return m;
}
inline float4x4 perspectiveMatrix(float fovy_deg, float aspect, float znear, float zfar) {
  const float ymax = znear * std::tan(fovy_deg * DEG_TO_RAD * 0.5f);
  const float xmax = ymax * aspect;
  float4x4 m;
  m.set_row(0, float4(znear / xmax, 0, 0, 0));
  m.set_row(1, float4(0, znear / ymax, 0, 0));
  m.set_row(2, float4(0, 0, -(zfar + znear) / (zfar - znear), -2.0f * zfar * znear / (zfar - znear)));
  m.set_row(3, float4(0, 0, -1, 0));
  return m;
}

// --- boxes -----------------------------------------------------------------
struct Box4f {
  float4 boxMin = float4(+MAXFLOAT, +MAXFLOAT, +MAXFLOAT, +MAXFLOAT);
  float4 boxMax = float4(-MAXFLOAT, -MAXFLOAT, -MAXFLOAT, -MAXFLOAT);
  void include(float4 p) {
    boxMin = min(boxMin, p);
    boxMax = max(boxMax, p);
  }
  void include(const Box4f& b) {
    boxMin = min(boxMin, b.boxMin);
    boxMax = max(boxMax, b.boxMax);
  }
  float4 center() const { return (boxMin + boxMax) * 0.5f; }
};

struct BBox3f {
  float3 boxMin = float3(+MAXFLOAT, +MAXFLOAT, +MAXFLOAT);
  float3 boxMax = float3(-MAXFLOAT, -MAXFLOAT, -MAXFLOAT);
  void include(float3 p) {
    boxMin = min(boxMin, p);
    boxMax = max(boxMax, p);
  }
  float3 center() const { return (boxMin + boxMax) * 0.5f; }
};

}  // namespace LiteMath
