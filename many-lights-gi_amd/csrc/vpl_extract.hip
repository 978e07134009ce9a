// VPL extraction: regular RSM sampling into VPL records.
// Follows gi/vpl_processor.comp:41-72 and VPLProcessor.cpp:73-101.
// Compiled with -ffp-contract=off: bit-exact against the FP32 oracle twin.
//
// One lane per VPL. Each lane reads its RSM texel (RGB diffuse, encoded
// normal, depth), unprojects with inverse(bias*P*V) using an explicit fused
// multiply-add chain, quantises the normal through the 10:10:10 packing, and
// also stores the lookAtRH basis so the splat and gather never recompute it,
// and the flux rounded to halves for the mixed-form gather.
#include "batch_device.hpp"
#include "kernels.hpp"
#include "mlgi_device.hpp"

namespace mlgi {
namespace {

using dev::f3;

// VPL i of a frame: RSM texel fetch, unprojection, normal quantisation,
// basis, flux; record written at shuffle[i].
__device__ __forceinline__ void vpl_extract_one(const float* __restrict__ diffuse,
                                                const float* __restrict__ normal,
                                                const float* __restrict__ depth, int w, int h,
                                                const Mat4& m, float light_intensity,
                                                const int32_t* __restrict__ shuffle, uint32_t i,
                                                uint32_t axis_x, uint32_t axis_y,
                                                float* __restrict__ records) {
  // uvec2(resultIndex % x, resultIndex / x) * samplerSize / lightsOnAxis
  const uint32_t tx = (i % axis_x) * (uint32_t)w / axis_x;
  const uint32_t ty = (i / axis_x) * (uint32_t)h / axis_y;
  const float uvx = (float)tx / (float)w;
  const float uvy = (float)ty / (float)h;
  float dr = 0.f, dg = 0.f, db = 0.f, nx = 0.f, ny = 0.f, nz = 0.f, d = 0.f;
  if (tx < (uint32_t)w && ty < (uint32_t)h) {  // out-of-range texelFetch reads 0
    const size_t p = (size_t)ty * (uint32_t)w + tx;
    dr = diffuse[p * 3 + 0];
    dg = diffuse[p * 3 + 1];
    db = diffuse[p * 3 + 2];
    nx = normal[p * 3 + 0];
    ny = normal[p * 3 + 1];
    nz = normal[p * 3 + 2];
    d = depth[p];
  }
  nx = nx * 2.0f - 1.0f;
  ny = ny * 2.0f - 1.0f;
  nz = nz * 2.0f - 1.0f;
  float r[4];
#pragma unroll
  for (int row = 0; row < 4; ++row) {
    // depth column first: it cancels against the constant column for a
    // perspective light, and one fused operation keeps that exact
    float acc = __builtin_fmaf(m.m[2 * 4 + row], d, m.m[3 * 4 + row]);
    acc = __builtin_fmaf(m.m[1 * 4 + row], uvy, acc);
    acc = __builtin_fmaf(m.m[0 * 4 + row], uvx, acc);
    r[row] = acc;
  }
  const uint32_t dst = shuffle ? (uint32_t)shuffle[i] : i;
  const uint32_t pk = dev::pack3sn(nx, ny, nz);
  const f3 un = dev::unpack3sn(pk);
  f3 s, u;
  dev::look_at_rh(un, s, u);
  const float px = r[0] / r[3], py = r[1] / r[3], pz = r[2] / r[3];
  const float plane = un.x * px + un.y * py + un.z * pz;  // n . position
  const float fr = dr * light_intensity, fg = dg * light_intensity, fb = db * light_intensity;
  // the flux as halves (round to nearest even) for the mixed-form gather
  const uint32_t hrg = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)fr) |
                       ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)fg) << 16);
  const uint32_t hb = __builtin_bit_cast(uint16_t, (_Float16)fb);
  // zero-flux flags for the gather's culling (bit 0: the floats are all
  // +-0; bit 1: the halves are all zero), beside the normal it reads anyway
  const uint32_t zero_flags =
      (((__float_as_uint(fr) | __float_as_uint(fg) | __float_as_uint(fb)) & 0x7FFFFFFFu) == 0u ? 1u : 0u) |
      ((hrg | hb) == 0u ? 2u : 0u);
  float4* out = reinterpret_cast<float4*>(records + (size_t)dst * kVplFloats);
  const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
  out[0] = make_float4(px, py, pz, plane);
  out[1] = make_float4(un.x, un.y, un.z, __uint_as_float(zero_flags));
  out[2] = make_float4(s.x, s.y, s.z, __uint_as_float(hrg));
  out[3] = make_float4(u.x, u.y, u.z, __uint_as_float(hb));
  out[4] = make_float4(fr, fg, fb, __uint_as_float(pk));
  out[5] = zero;
  out[6] = zero;
  out[7] = zero;
}

__global__ __launch_bounds__(64) void vpl_extract_kernel(
    const float* __restrict__ diffuse, const float* __restrict__ normal,
    const float* __restrict__ depth, int w, int h, Mat4 m, float light_intensity,
    const int32_t* __restrict__ shuffle, int n, uint32_t axis_x, uint32_t axis_y,
    float* __restrict__ records) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint32_t)n) return;
  vpl_extract_one(diffuse, normal, depth, w, h, m, light_intensity, shuffle, i, axis_x, axis_y, records);
}

// Batched: one lane per VPL of any frame.
__global__ __launch_bounds__(256) void batch_vpl_kernel(const BatchFrame* __restrict__ frames, int n,
                                                        uint32_t total) {
  for (uint32_t item = blockIdx.x * blockDim.x + threadIdx.x; item < total;
       item += gridDim.x * blockDim.x) {
    const BatchFrame& f = frames[batch_find<&BatchFrame::first_vpl>(frames, n, item)];
    vpl_extract_one(f.rsm_diffuse, f.rsm_normal, f.rsm_depth, f.rsm_w, f.rsm_h, f.inv_bvp,
                    f.light_intensity, f.shuffle, item - f.first_vpl, f.axis_x, f.axis_y,
                    const_cast<float*>(f.g.vpl));
  }
}

}  // namespace

hipError_t launch_vpl_extract(hipStream_t s, const float* diffuse, const float* normal,
                              const float* depth, int w, int h, const Mat4& inv_bvp,
                              float light_intensity, const int32_t* shuffle, int n,
                              float* records) {
  uint32_t s1d = 0;
  while ((uint64_t)s1d * s1d < (uint64_t)n) ++s1d;  // uint(ceil(sqrt(totalVplCount)))
  const uint32_t axis_y = (uint32_t)n / s1d;          // totalWorkCount / rsmSamples1d
  const int blocks = (n + 63) / 64;
  hipLaunchKernelGGL(vpl_extract_kernel, dim3(blocks), dim3(64), 0, s, diffuse, normal, depth, w,
                     h, inv_bvp, light_intensity, shuffle, n, s1d, axis_y, records);
  return hipGetLastError();
}

hipError_t launch_batch_vpl(hipStream_t s, const BatchFrame* frames, int n, uint32_t total) {
  if (total == 0) return hipSuccess;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL(batch_vpl_kernel, dim3(blocks), dim3(256), 0, s, frames, n, total);
  return hipGetLastError();
}

}  // namespace mlgi
