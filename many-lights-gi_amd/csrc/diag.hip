// Measurement aid: the FP32 vector FMA rate of the device, the peak that the
// VALU-bound kernels (pull-push, final gather) are compared against.
//
// Every lane of a 2048 x 256 grid (8 workgroups per CU) runs kChains
// independent fused multiply-add chains; with kChains in flight per lane and
// 8 waves per SIMD, the FMA latency is hidden and the loop issues one VALU
// instruction per chain step (the loop control is scalar).
#include "kernels.hpp"

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

}  // namespace

uint64_t valu_probe_fmas(int blocks, int iters) { return (uint64_t)blocks * 256u * (uint64_t)iters * kChains; }

hipError_t launch_valu_probe(hipStream_t s, float* out, int blocks, int iters) {
  hipLaunchKernelGGL(valu_probe_kernel, dim3(blocks), dim3(256), 0, s, out, iters, 0.999999f, 1e-7f);
  return hipGetLastError();
}

}  // namespace mlgi
