// Device helpers shared by the gfx950 kernels: packing formats and the
// paraboloid ISM geometry with a fixed IEEE evaluation order.
//
// The translation units that include this for bit-exact stages (VPL
// extraction, splat, pull-push) are compiled with -ffp-contract=off, so every
// product and sum below rounds separately; fused multiply-adds appear only
// where __builtin_fmaf is written out.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlgi {
namespace dev {

struct f3 {
  float x, y, z;
};

__device__ __forceinline__ f3 make(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
// dot accumulates x, y, z left to right (GLSL dot / mat3*vec3 column order)
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// c / 255 for a byte c (unpackUnorm4x8), correctly rounded: q = RN(c y)
// with y = RN(1/255), the exact FMA remainder r = c - 255 q, then RN(q + r y)
// (Markstein's form; equal to the IEEE quotient for all 256 bytes, which the
// CPU tests check exhaustively). Three operations instead of the scaled
// division sequence.
__device__ __forceinline__ float unorm8(uint32_t c) {
  const float x = (float)c, y = 1.0f / 255.0f;
  const float q = x * y;
  return __builtin_fmaf(__builtin_fmaf(-q, 255.0f, x), y, q);
}

// RN(1 / b) for a normal float b: the hardware reciprocal (within one ulp)
// refined by one Newton step, y + y (1 - b y) in two FMAs. Equal to the IEEE
// quotient 1 / b for every significand (mlgi_check_division runs that check
// on the device: all 2^23 significands at eleven exponents from 2^-100 to
// 2^100). A zero, infinite or NaN b gives NaN.
__device__ __forceinline__ float recip_rn(float b) {
  const float y = __builtin_amdgcn_rcpf(b);
  return __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
}

// RN(a / b) from y = RN(1 / b): q = RN(a y) lies within an ulp of a / b, the
// remainder a - q b is exact under an FMA, and RN(q + r y) is the correctly
// rounded quotient (Markstein's theorem), for quotients and remainders
// clear of overflow and underflow. Three operations; the reciprocal is
// shared by every quotient of one divisor. An exact zero quotient may
// carry the other sign than the IEEE one.
__device__ __forceinline__ float div_rn(float a, float b, float y) {
  const float q = a * y;
  return __builtin_fmaf(__builtin_fmaf(-q, b, a), y, q);
}

// GLSL uint(float): truncation, negative values saturate to 0.
__device__ __forceinline__ uint32_t to_uint(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)v;
}

// pack3SNToFloat / unpack3SNFromFloat (floatpacking.glsl:26-66); the packed
// value is kept as uint32 so that no float path can flush its denormal bit
// patterns.
__device__ __forceinline__ uint32_t pack3un(float x, float y, float z) {
  uint32_t u = to_uint(x * 1023.0f + 0.5f);
  u |= to_uint(y * 1023.0f + 0.5f) << 10;
  u |= to_uint(z * 1021.0f + 1.5f) << 20;
  return u;
}
__device__ __forceinline__ uint32_t pack3sn(float x, float y, float z) {
  return pack3un(x * 0.5f + 0.5f, y * 0.5f + 0.5f, z * 0.5f + 0.5f);
}
__device__ __forceinline__ f3 unpack3sn(uint32_t u) {
  float a = (float)(u & 0x3FFu) / 1023.0f;
  float b = (float)((u >> 10) & 0x3FFu) / 1023.0f;
  float c = ((float)((u >> 20) & 0x3FFu) - 1.5f) / 1021.0f;
  return f3{a * 2.0f - 1.0f, b * 2.0f - 1.0f, c * 2.0f - 1.0f};
}

// lookAtRH (ism_utils.glsl:5-13): rows s, u, -f of the VPL view basis.
__device__ __forceinline__ void look_at_rh(f3 f, f3& s, f3& u) {
  const f3 up = f3{0.0f, 1.0f, 0.01f};
  f3 c = cross(f, up);
  float l = sqrtf(dot(c, c));
  s = f3{c.x / l, c.y / l, c.z / l};
  u = cross(s, f);
}

// packHalf2x16 / unpackHalf2x16 with round-to-nearest-even conversion.
__device__ __forceinline__ uint32_t pack_half2(float x, float y) {
  _Float16 hx = (_Float16)x, hy = (_Float16)y;
  return (uint32_t)__builtin_bit_cast(uint16_t, hx) |
         ((uint32_t)__builtin_bit_cast(uint16_t, hy) << 16);
}
__device__ __forceinline__ void unpack_half2(uint32_t u, float& x, float& y) {
  x = (float)__builtin_bit_cast(_Float16, (uint16_t)(u & 0xFFFFu));
  y = (float)__builtin_bit_cast(_Float16, (uint16_t)(u >> 16));
}

}  // namespace dev
}  // namespace mlgi
