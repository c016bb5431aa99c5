// Internal helpers shared by the lgx HIP kernels and host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "lgx.h"

#define LGX_HD __host__ __device__ __forceinline__
#define LGX_D __device__ __forceinline__

namespace lgx {

LGX_HD int format_bytes(int f) {
  switch (f) {
    case LGX_R32F: return 4;
    case LGX_RG32F: return 8;
    case LGX_RGBA32F: return 16;
    case LGX_RG16F: return 4;
    case LGX_RGBA16F: return 8;
    case LGX_RGBA8: return 4;
    case LGX_RGBA64F: return 32;
    case LGX_R16F: return 2;
    default: return 0;
  }
}
LGX_HD int format_channels(int f) {
  switch (f) {
    case LGX_R32F: case LGX_R16F: return 1;
    case LGX_RG32F: case LGX_RG16F: return 2;
    default: return 4;
  }
}

inline bool image_ok(const lgx_image *im) {
  return im && im->ptr && im->w > 0 && im->h > 0 && im->pitch >= im->w && format_bytes(im->format) > 0 &&
         im->reserved == 0;
}

// IEEE binary16 <-> binary32, round to nearest even (host path; the device
// path uses the hardware conversion, which rounds the same way).
inline uint16_t host_f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? (0x200u | ((ax >> 13) & 0x3ffu)) : 0u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
  if (ax < 0x38800000u) {
    uint32_t e = ax >> 23;
    if (e < 102) return (uint16_t)sign;
    uint32_t m = (ax & 0x7fffffu) | 0x800000u, sh = 126 - e;
    uint32_t q = m >> sh, rem = m & ((1u << sh) - 1), half = 1u << (sh - 1);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  return (uint16_t)(sign | ((ax + 0xfffu + ((ax >> 13) & 1u) - 0x38000000u) >> 13));
}
inline float host_f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu, x;
  if (e == 0) {
    float v = (float)m * (1.0f / 16777216.0f);
    memcpy(&x, &v, 4);
    x |= sign;
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &x, 4);
  return f;
}

LGX_HD float h2f(uint16_t h) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (float)__builtin_bit_cast(_Float16, h);
#else
  return host_f16_to_f32(h);
#endif
}
LGX_HD uint16_t f2h(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_bit_cast(uint16_t, (_Float16)f);
#else
  return host_f32_to_f16(f);
#endif
}

// Texel loads, widened to float.  FMT is an LGX_* format.
template <int FMT> struct Texel;
template <> struct Texel<LGX_R32F> {
  static LGX_D float4 load(const void *base, size_t i) {
    float v = ((const float *)base)[i];
    return make_float4(v, 0.f, 0.f, 1.f);
  }
};
template <> struct Texel<LGX_RG32F> {
  static LGX_D float4 load(const void *base, size_t i) {
    float2 v = ((const float2 *)base)[i];
    return make_float4(v.x, v.y, 0.f, 1.f);
  }
};
template <> struct Texel<LGX_RGBA32F> {
  static LGX_D float4 load(const void *base, size_t i) { return ((const float4 *)base)[i]; }
};
template <> struct Texel<LGX_RG16F> {
  static LGX_D float4 load(const void *base, size_t i) {
    uint32_t v = ((const uint32_t *)base)[i];
    return make_float4(h2f((uint16_t)(v & 0xffffu)), h2f((uint16_t)(v >> 16)), 0.f, 1.f);
  }
};
template <> struct Texel<LGX_RGBA16F> {
  static LGX_D float4 load(const void *base, size_t i) {
    uint2 v = ((const uint2 *)base)[i];
    return make_float4(h2f((uint16_t)(v.x & 0xffffu)), h2f((uint16_t)(v.x >> 16)), h2f((uint16_t)(v.y & 0xffffu)),
                       h2f((uint16_t)(v.y >> 16)));
  }
};
template <> struct Texel<LGX_RGBA8> {
  static LGX_D float4 load(const void *base, size_t i) {
    uint32_t v = ((const uint32_t *)base)[i];
    const float s = 1.0f / 255.0f;
    return make_float4((float)(v & 0xffu) * s, (float)((v >> 8) & 0xffu) * s, (float)((v >> 16) & 0xffu) * s,
                       (float)(v >> 24) * s);
  }
};

// Stores with rounding to the storage format.
template <int FMT> struct Store;
template <> struct Store<LGX_R32F> {
  static LGX_D void store(void *base, size_t i, float4 v) { ((float *)base)[i] = v.x; }
  static LGX_D float4 round(float4 v) { return v; }
};
template <> struct Store<LGX_RG32F> {
  static LGX_D void store(void *base, size_t i, float4 v) { ((float2 *)base)[i] = make_float2(v.x, v.y); }
  static LGX_D float4 round(float4 v) { return v; }
};
template <> struct Store<LGX_RGBA32F> {
  static LGX_D void store(void *base, size_t i, float4 v) { ((float4 *)base)[i] = v; }
  static LGX_D float4 round(float4 v) { return v; }
};
template <> struct Store<LGX_RG16F> {
  static LGX_D void store(void *base, size_t i, float4 v) {
    ((uint32_t *)base)[i] = (uint32_t)f2h(v.x) | ((uint32_t)f2h(v.y) << 16);
  }
  static LGX_D float4 round(float4 v) { return make_float4(h2f(f2h(v.x)), h2f(f2h(v.y)), v.z, v.w); }
};
template <> struct Store<LGX_RGBA16F> {
  static LGX_D void store(void *base, size_t i, float4 v) {
    ((uint2 *)base)[i] = make_uint2((uint32_t)f2h(v.x) | ((uint32_t)f2h(v.y) << 16),
                                    (uint32_t)f2h(v.z) | ((uint32_t)f2h(v.w) << 16));
  }
  static LGX_D float4 round(float4 v) {
    return make_float4(h2f(f2h(v.x)), h2f(f2h(v.y)), h2f(f2h(v.z)), h2f(f2h(v.w)));
  }
};

// Octahedral unit-vector decode (Cigolle et al. 2014, "A Survey of
// Efficient Representations for Independent Unit Vectors").
template <class T> struct Vec3 {
  T x, y, z;
};
template <class T> LGX_HD Vec3<T> v3(T x, T y, T z) { return Vec3<T>{x, y, z}; }
template <class T> LGX_HD Vec3<T> operator+(Vec3<T> a, Vec3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> LGX_HD Vec3<T> operator-(Vec3<T> a, Vec3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> LGX_HD Vec3<T> operator*(Vec3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> LGX_HD T dot3(Vec3<T> a, Vec3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> LGX_HD Vec3<T> cross3(Vec3<T> a, Vec3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> LGX_HD Vec3<T> normalize3(Vec3<T> a) {
  T len = sqrt(dot3(a, a));
  return {a.x / len, a.y / len, a.z / len};
}
template <class T> LGX_HD Vec3<T> oct_decode(T ex, T ey) {
  Vec3<T> n{ex, ey, T(1) - fabs(ex) - fabs(ey)};
  if (n.z < T(0)) {
    T ox = (T(1) - fabs(n.y)) * (n.x >= T(0) ? T(1) : T(-1));
    T oy = (T(1) - fabs(n.x)) * (n.y >= T(0) ? T(1) : T(-1));
    n.x = ox;
    n.y = oy;
  }
  return normalize3(n);
}
template <class T> LGX_HD void oct_encode(Vec3<T> n, T &ex, T &ey) {
  T s = fabs(n.x) + fabs(n.y) + fabs(n.z);
  T px = n.x / s, py = n.y / s;
  if (n.z < T(0)) {
    T ox = (T(1) - fabs(py)) * (px >= T(0) ? T(1) : T(-1));
    T oy = (T(1) - fabs(px)) * (py >= T(0) ? T(1) : T(-1));
    px = ox;
    py = oy;
  }
  ex = px;
  ey = py;
}

// Column-major 4x4 (glm layout: m[col * 4 + row]).
struct Mat4 {
  float m[16];
};
Mat4 mat4_mul(const Mat4 &a, const Mat4 &b);
Mat4 mat4_inverse(const Mat4 &a);
struct Mat4d {
  double m[16];
};
Mat4d mat4d_inverse(const Mat4d &a);
Mat4d mat4d_mul(const Mat4d &a, const Mat4d &b);

int hip_status(hipError_t e);

// The gather launch shape with the same tile, registers and step loop that
// reads the pyramid levels as they are (tile itself if it does).
int gather_unswizzled_tile(int tile);

// XCD-aware block remap: MI355X dispatches consecutive workgroups round
// robin over its 8 XCDs (each with its own L2).  Remap so XCD x works on the
// x-th contiguous eighth of the logical block range, keeping neighbouring
// tiles (which share pyramid texels) in one L2.
LGX_HD uint32_t xcd_remap(uint32_t b, uint32_t n) {
  const uint32_t X = 8;
  uint32_t per = n / X, rem = n % X;  // XCD x gets per (+1 if x < rem) blocks
  uint32_t x = b % X, i = b / X;
  // XCD x receives blocks x, x+8, ...: per + (x < rem) of them, so i is
  // always below that count and the map is a bijection of [0, n).
  uint32_t start = x * per + (x < rem ? x : rem);
  return start + i;
}

}  // namespace lgx
