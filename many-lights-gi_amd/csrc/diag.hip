// Measurement aid: the FP32 vector FMA rate of the device, the peak that the
// VALU-bound kernels (pull-push, final gather) are compared against. And a
// check of the device arithmetic the splat's bit-exactness rests on: the
// reciprocal and quotient forms of mlgi_device.hpp against IEEE division.
//
// Every lane of a 2048 x 256 grid (8 workgroups per CU) runs kChains
// independent fused multiply-add chains; with kChains in flight per lane and
// 8 waves per SIMD, the FMA latency is hidden and the loop issues one VALU
// instruction per chain step (the loop control is scalar).
#include "kernels.hpp"
#include "mlgi_device.hpp"

namespace mlgi {
namespace {

constexpr int kChains = 16;

__global__ __launch_bounds__(256) void valu_probe_kernel(float* __restrict__ out, int iters, float a, float b) {
  float x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = (float)(threadIdx.x + c) * 1e-3f;
#pragma unroll 4
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = __builtin_fmaf(x[c], a, b);
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  if (s == -1.0f) out[blockIdx.x] = s;  // never true; keeps the chains live
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// Thread m (< 2^23) takes significand m at the exponents of kDivExps
// (2^-100 .. 2^100 in steps of 2^20): the refined reciprocal against the IEEE
// 1 / b, and kDivNums quotients a / b for numerators of random significand,
// sign and scale (2^-40 .. 2^40 times that of b, within 2^-100 .. 2^100, so
// that quotient and remainder stay normal floats, where the correction step
// is exact) against the IEEE a / b. counts: reciprocal and quotient
// mismatches.
constexpr int kNumDivExps = 11;
constexpr int kDivExps[kNumDivExps] = {127 - 100, 127 - 80, 127 - 60, 127 - 40, 127 - 20, 127,
                                       127 + 20,  127 + 40, 127 + 60, 127 + 80, 127 + 100};
constexpr int kDivNums = 4;

__global__ __launch_bounds__(256) void division_check_kernel(uint32_t* __restrict__ counts, uint32_t seed) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= (1u << 23)) return;
  uint32_t bad_r = 0, bad_q = 0;
#pragma unroll
  for (int k = 0; k < kNumDivExps; ++k) {
    const float b = __uint_as_float(((uint32_t)kDivExps[k] << 23) | m);
    const float y = dev::recip_rn(b);
    if (__float_as_uint(y) != __float_as_uint(1.0f / b)) ++bad_r;
#pragma unroll
    for (int j = 0; j < kDivNums; ++j) {
      const uint32_t h = mix32(m * 32u + (uint32_t)(k * kDivNums + j) + seed);
      const int e = min(127 + 100, max(127 - 100, kDivExps[k] - 40 + (int)(h % 81u)));
      const float a = __uint_as_float((h & 0x80000000u) | ((uint32_t)e << 23) | (mix32(h) & 0x7FFFFFu));
      if (__float_as_uint(dev::div_rn(a, b, y)) != __float_as_uint(a / b)) ++bad_q;
    }
  }
  if (bad_r) atomicAdd(&counts[0], bad_r);
  if (bad_q) atomicAdd(&counts[1], bad_q);
}

}  // namespace

uint64_t division_check_quotients() { return (uint64_t)(1u << 23) * kNumDivExps * kDivNums; }

hipError_t launch_division_check(hipStream_t s, uint32_t* counts, uint32_t seed) {
  hipLaunchKernelGGL(division_check_kernel, dim3((1u << 23) / 256u), dim3(256), 0, s, counts, seed);
  return hipGetLastError();
}

uint64_t valu_probe_fmas(int blocks, int iters) { return (uint64_t)blocks * 256u * (uint64_t)iters * kChains; }

hipError_t launch_valu_probe(hipStream_t s, float* out, int blocks, int iters) {
  hipLaunchKernelGGL(valu_probe_kernel, dim3(blocks), dim3(256), 0, s, out, iters, 0.999999f, 1e-7f);
  return hipGetLastError();
}

}  // namespace mlgi
