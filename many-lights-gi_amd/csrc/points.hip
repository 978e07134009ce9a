// Device point generation: the area-proportional point set of a scene
// (what the reference's tessellation and geometry shaders emit per frame,
// ism.tesc:11-23, ism.geom:49-69) sampled on the GPU and grouped by bin,
// for scenes whose point sets are built at run time. Each lane samples one
// point with the host sampler's double-precision code (scene_trace.hpp,
// compiled with -ffp-contract=off in both), so record i and its bin equal
// the host's; bins are counted with atomics, scanned by one workgroup and
// filled by an atomic cursor per bin. Within a bin the records come in no
// fixed order: every consumer (the splat) takes a minimum over the bin's
// splats, which no order changes.
#include "kernels.hpp"
#include "scene_trace.hpp"

namespace mlgi {
namespace {

using namespace mlgi_scene_impl;

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* __restrict__ p, int n, uint32_t v) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = v;
}

__global__ __launch_bounds__(256) void point_sample_kernel(const Tri* __restrict__ tris,
                                                           const double* __restrict__ cdf, int nt, double total,
                                                           uint64_t seed, float r_packed, int n_bins, int n,
                                                           float4* __restrict__ tmp, int* __restrict__ bin,
                                                           uint32_t* __restrict__ count) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float rec[4];
  int b;
  sample_point(tris, cdf, nt, total, seed, (uint64_t)i, r_packed, n_bins, rec, &b);
  tmp[i] = make_float4(rec[0], rec[1], rec[2], rec[3]);
  bin[i] = b;
  __hip_atomic_fetch_add(&count[b], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of count[0, n_bins) into offsets[0, n_bins] (one workgroup
// of 1024 lanes, each owning a contiguous run of bins); cursor = offsets.
__global__ __launch_bounds__(1024) void bin_scan_kernel(const uint32_t* __restrict__ count, int n_bins,
                                                        uint32_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ cursor) {
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x;
  const int per = (n_bins + 1023) / 1024;
  const int b0 = min(n_bins, t * per), b1 = min(n_bins, b0 + per);
  uint32_t s = 0;
  for (int b = b0; b < b1; ++b) s += count[b];
  part[t] = s;
  __syncthreads();
  // inclusive Hillis-Steele scan of the 1024 run sums
  for (int o = 1; o < 1024; o <<= 1) {
    const uint32_t v = t >= o ? part[t - o] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = t > 0 ? part[t - 1] : 0u;
  for (int b = b0; b < b1; ++b) {
    offsets[b] = run;
    cursor[b] = run;
    run += count[b];
  }
  if (t == 1023) offsets[n_bins] = part[1023];
}

__global__ __launch_bounds__(256) void point_scatter_kernel(const float4* __restrict__ tmp,
                                                            const int* __restrict__ bin, int n,
                                                            uint32_t* __restrict__ cursor,
                                                            float4* __restrict__ records) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t dst = __hip_atomic_fetch_add(&cursor[bin[i]], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  records[dst] = tmp[i];
}

unsigned blocks_for(int n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t launch_sample_points(hipStream_t s, const Tri* tris, const double* cdf, int nt, double total,
                                uint64_t seed, float r_packed, int n_bins, int n, float* tmp, int* bin,
                                uint32_t* count, uint32_t* cursor, float* records, uint32_t* offsets) {
  hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(n_bins) < 1024 ? blocks_for(n_bins) : 1024), dim3(256), 0, s,
                     count, n_bins, 0u);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (n > 0) {
    hipLaunchKernelGGL(point_sample_kernel, dim3(blocks_for(n)), dim3(256), 0, s, tris, cdf, nt, total, seed, r_packed,
                       n_bins, n, reinterpret_cast<float4*>(tmp), bin, count);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(1024), 0, s, count, n_bins, offsets, cursor);
  if ((e = hipGetLastError()) != hipSuccess || n == 0) return e;
  hipLaunchKernelGGL(point_scatter_kernel, dim3(blocks_for(n)), dim3(256), 0, s, reinterpret_cast<const float4*>(tmp),
                     bin, n, cursor, reinterpret_cast<float4*>(records));
  return hipGetLastError();
}

}  // namespace mlgi
