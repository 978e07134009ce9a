// Deferred composite: direct light with a VSM shadow, GI ambient, Reinhard
// tone mapping and gamma 2.2 (deferredshading.frag:47-81 with
// shadowmapping.glsl:6-34 and srgb_utils.glsl, bound by
// DeferredShadingStage::process, DeferredShadingStage.cpp:76-131), one lane
// per pixel over 16 x 16 tiles. Elementwise and bandwidth-bound: 44 bytes
// read (depth, normal, diffuse, specular, GI, occlusion) and 12 written per
// pixel, plus four bilinear taps of the VSM moments.
//
// Conventions (the oracle's): the pixel is sampled at its centre; the world
// position is inverse(P*V) of the centre's NDC; the VSM is sampled
// bilinearly with a (0, 0) border; pow(0, y) = 0; a null specular buffer
// reads 0 and a null occlusion buffer reads 1. FP32 throughout.
#include "kernels.hpp"

namespace mlgi {
namespace {

__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float smoothstep_f(float e0, float e1, float x) {
  const float t = clamp01((x - e0) / (e1 - e0));
  return t * t * (3.0f - 2.0f * t);
}
__device__ __forceinline__ float pow_pos(float x, float y) { return x > 0.0f ? powf(x, y) : 0.0f; }

// shadowmapping.glsl:13-20
__device__ __forceinline__ float vsm_factor(float mx, float my, float compare) {
  const float p = smoothstep_f(compare - 0.001f, compare, mx);
  // E[d^2] - E[d]^2 cancels: one rounding (fused) instead of two
  const float variance = fmaxf(__builtin_fmaf(-mx, mx, my), 0.001f);
  const float d = compare - mx;
  const float p_max = clamp01((variance / (variance + d * d) - 0.1f) / (1.0f - 0.1f));
  return clamp01(fmaxf(p, p_max));
}

__global__ __launch_bounds__(256) void composite_kernel(CompositeArgs a) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= a.w || y >= a.h) return;
  const size_t c = (size_t)y * a.w + x;
  const float u = ((float)x + 0.5f) / (float)a.w, v = ((float)y + 0.5f) / (float)a.h;
  const float ndc[4] = {u * 2.0f - 1.0f, v * 2.0f - 1.0f, a.depth[c] * 2.0f - 1.0f, 1.0f};
  float r[4];
#pragma unroll
  for (int row = 0; row < 4; ++row) {
    float acc = 0.0f;
#pragma unroll
    for (int col = 0; col < 4; ++col) acc += a.vp_inv[col * 4 + row] * ndc[col];
    r[row] = acc;
  }
  const float wx = r[0] / r[3], wy = r[1] / r[3], wz = r[2] / r[3];
  const float nx = a.normal[c * 3] * 2.0f - 1.0f, ny = a.normal[c * 3 + 1] * 2.0f - 1.0f,
              nz = a.normal[c * 3 + 2] * 2.0f - 1.0f;
  const float lx = a.inv_light_dir[0], ly = a.inv_light_dir[1], lz = a.inv_light_dir[2];
  float vx = a.eye[0] - wx, vy = a.eye[1] - wy, vz = a.eye[2] - wz;
  const float vl = sqrtf(vx * vx + vy * vy + vz * vz);
  vx /= vl, vy /= vl, vz /= vl;
  float hx = lx + vx, hy = ly + vy, hz = lz + vz;
  const float hl = sqrtf(hx * hx + hy * hy + hz * hz);
  hx /= hl, hy /= hl, hz /= hl;
  const float ndotl = nx * lx + ny * ly + nz * lz;
  const float ndoth = nx * hx + ny * hy + nz * hz;
  float sc[4];
  const float wp[4] = {wx, wy, wz, 1.0f};
#pragma unroll
  for (int row = 0; row < 4; ++row) {
    float acc = 0.0f;
#pragma unroll
    for (int col = 0; col < 4; ++col) acc += a.blvp[col * 4 + row] * wp[col];
    sc[row] = acc;
  }
  // bilinear VSM moments, border (0, 0)
  const float s = sc[0] / sc[3] * (float)a.vw - 0.5f, t = sc[1] / sc[3] * (float)a.vh - 0.5f;
  const float fs = floorf(s), ft = floorf(t);
  const float fa = s - fs, fb = t - ft;
  float mx = 0.0f, my = 0.0f;
  if (fabsf(fs) < 1e9f && fabsf(ft) < 1e9f) {  // far outside: all border
    const int i0 = (int)fs, j0 = (int)ft;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = i0 + (k & 1), j = j0 + (k >> 1);
      const float wgt = ((k & 1) ? fa : 1.0f - fa) * ((k >> 1) ? fb : 1.0f - fb);
      if (i < 0 || j < 0 || i >= a.vw || j >= a.vh) continue;
      const float2 m = reinterpret_cast<const float2*>(a.vsm)[(size_t)j * a.vw + i];
      mx += wgt * m.x;
      my += wgt * m.y;
    }
  }
  const float dlx = wx - a.light_pos[0], dly = wy - a.light_pos[1], dlz = wz - a.light_pos[2];
  const float dist = sqrtf(dlx * dlx + dly * dly + dlz * dlz);
  const float shadow = vsm_factor(mx, my, dist) * (sc[3] > 0.0f ? 1.0f : 0.0f);
  const float occ = a.occlusion ? a.occlusion[c] : 1.0f;
  const float spec_f = pow_pos(fmaxf(0.0f, ndoth), 20.0f);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float dif = pow_pos(a.diffuse[c * 3 + k], 2.2f);
    const float spc = a.specular ? pow_pos(a.specular[c * 3 + k], 2.2f) : 0.0f;
    const float ambient = a.gi[c * 3 + k] * dif * occ;
    const float dterm = dif * (fmaxf(0.0f, ndotl) * shadow) * a.light_intensity;
    const float sterm = 0.75f * spc * spec_f * shadow;
    float col = (ambient + dterm + sterm) * a.exposure;
    col = col / (1.0f + col);
    a.out[c * 3 + k] = pow_pos(col, 1.0f / 2.2f);
  }
}

}  // namespace

hipError_t launch_composite(hipStream_t s, const CompositeArgs& a) {
  const dim3 grid((a.w + 15) / 16, (a.h + 15) / 16);
  hipLaunchKernelGGL(composite_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace mlgi
