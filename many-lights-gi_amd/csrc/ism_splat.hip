// ISM point splatting with 32-bit atomicMin on packed depth.
// Follows ism/ism.comp:54-139 (usePushPull path) and the atlas clear of
// ImperfectShadowmap.cpp:212. Compiled with -ffp-contract=off: the atlas is
// bit-exact against the FP32 oracle twin (integer atomicMin is order-free).
//
// One workgroup per point bin b (256 lanes by default, so a bin of a few
// hundred points is one pass and every SIMD holds enough waves to cover the
// latency of the point loads and atomics). The 16 candidate VPLs (b+i) mod N
// are staged in LDS once per workgroup (1 KiB); each lane owns one point per
// iteration, tests all 16 candidates from broadcast LDS reads (uniform
// addresses, no bank conflicts), keeps the first four that pass both
// hemisphere tests as a bit mask, and projects only those. The projected
// texel receives a no-return, relaxed, agent-scope atomic umin, which the L2
// executes without a round trip to the wave.
#include "batch_device.hpp"
#include "kernels.hpp"
#include "mlgi_device.hpp"

namespace mlgi {
namespace {

using dev::f3;

constexpr int kMaxTest = 16;    // maxVplTestCount (ism.comp:44)
constexpr int kMaxCollect = 4;  // maxVplCollectCount (ism.comp:49)

__global__ __launch_bounds__(256) void atlas_clear_kernel(uint4* __restrict__ atlas, size_t n4) {
  const uint4 ones = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (size_t)gridDim.x * blockDim.x)
    atlas[i] = ones;
}

struct Candidate {
  float px, py, pz;
  float nx, ny, nz;
  float sx, sy, sz;
  float ux, uy, uz;
  int ism;  // ISM index == global VPL id (scaleISMs == false)
  int tile_x, tile_y;
  int pad;
};

// Bin b of one frame: stage the 16 candidates in LDS, then every lane splats
// one point per pass. The first point record is requested before the staging
// barrier and each next one while the current one is processed.
template <int kThreads>
__device__ __forceinline__ void splat_bin(Candidate* cand, const float4* __restrict__ points,
                                          const uint32_t* __restrict__ bin_offsets, uint32_t n_points,
                                          const float* __restrict__ vpl, int n_vpl, int n1d,
                                          int atlas_side, float z_far, uint32_t* __restrict__ atlas,
                                          int b) {
  const int lane = threadIdx.x;
  // clamp to the point count so malformed offsets cannot index past the buffer
  const uint32_t end = min(bin_offsets[b + 1], n_points);
  const uint32_t begin = min(bin_offsets[b], end);
  uint32_t p = begin + lane;
  float4 rec = p < end ? points[p] : make_float4(0.f, 0.f, 0.f, 0.f);
  if (lane < kMaxTest) {
    const int idx = (b + lane) % n_vpl;  // index %= totalVplCount
    const float4* r = reinterpret_cast<const float4*>(vpl + (size_t)idx * kVplFloats);
    const float4 r0 = r[kVplPos / 4], r1 = r[kVplNormal / 4], r2 = r[kVplS / 4], r3 = r[kVplU / 4];
    Candidate c;
    c.px = r0.x; c.py = r0.y; c.pz = r0.z;
    c.nx = r1.x; c.ny = r1.y; c.nz = r1.z;
    c.sx = r2.x; c.sy = r2.y; c.sz = r2.z;
    c.ux = r3.x; c.uy = r3.y; c.uz = r3.z;
    c.ism = idx;
    // int(ismIndex / ismIndices1d), exact for idx < 2^24
    c.tile_y = idx / n1d;
    c.tile_x = idx - c.tile_y * n1d;
    c.pad = 0;
    cand[lane] = c;
  }
  __syncthreads();

  const float inv_n1d = 1.0f / (float)n1d;  // exact: n1d is a power of two
  const float side = (float)atlas_side;
  while (p < end) {
    const uint32_t pn = p + kThreads;
    const float4 next = pn < end ? points[pn] : rec;
    const f3 pos = dev::make(rec.x, rec.y, rec.z);
    const uint32_t bits = __float_as_uint(rec.w);
    // unpackUnorm4x8: c / 255
    const float ux = (float)(bits & 0xFFu) / 255.0f;
    const float uy = (float)((bits >> 8) & 0xFFu) / 255.0f;
    const float uz = (float)((bits >> 16) & 0xFFu) / 255.0f;
    const float uw = (float)(bits >> 24) / 255.0f;
    const f3 pn3 = dev::make(ux * 2.0f - 1.0f, uy * 2.0f - 1.0f, uz * 2.0f - 1.0f);
    const float radius = uw * 25.0f;
    const uint32_t radius_byte = dev::to_uint(radius * 10.0f / 2.0f);  // sqrt(4) == 2

    // gather up to kMaxCollect candidates that pass both hemisphere tests
    uint32_t keep = 0;
    int found = 0;
#pragma unroll
    for (int i = 0; i < kMaxTest; ++i) {
      const Candidate& c = cand[i];
      const f3 rel = dev::sub(pos, dev::make(c.px, c.py, c.pz));
      const f3 nrel = dev::make(-rel.x, -rel.y, -rel.z);
      const bool cull = dev::dot(dev::make(c.nx, c.ny, c.nz), rel) < 0.0f || dev::dot(pn3, nrel) < 0.0f;
      if (!cull && found < kMaxCollect) {
        keep |= 1u << i;
        ++found;
      }
    }

    while (keep) {
      const int i = __builtin_ctz(keep);
      keep &= keep - 1u;
      const Candidate& c = cand[i];
      const f3 rel = dev::sub(pos, dev::make(c.px, c.py, c.pz));
      const float dist = sqrtf(dev::dot(rel, rel));
      // paraboloid_project(rel, dist, n, zFar, idx, n1d, preserveSign = true)
      float vx = dev::dot(dev::make(c.sx, c.sy, c.sz), rel);
      float vy = dev::dot(dev::make(c.ux, c.uy, c.uz), rel);
      float vz = dev::dot(dev::make(-c.nx, -c.ny, -c.nz), rel);
      const float sign_v = vz > 0.0f ? 1.0f : (vz < 0.0f ? -1.0f : 0.0f);
      vx = vx / dist;
      vy = vy / dist;
      vz = vz / dist;
      vz = 1.0f - vz;
      vx = vx / vz;
      vy = vy / vz;
      vz = dist / z_far;
      vz = vz * -sign_v;
      vx = (vx + 1.0f) * 0.5f;
      vy = (vy + 1.0f) * 0.5f;
      vx = (vx + (float)c.tile_x) * inv_n1d;
      vy = (vy + (float)c.tile_y) * inv_n1d;
      vx = vx * side;
      vy = vy * side;
      vz = vz * 16777216.0f;
      const uint32_t value = (dev::to_uint(vz) << 8) | radius_byte;
      // ivec2(v.xy) truncates; imageAtomicMin outside the image is a no-op
      const float fx = truncf(vx), fy = truncf(vy);
      if (!(fx >= 0.0f && fx < side && fy >= 0.0f && fy < side)) continue;
      const size_t texel = (size_t)(uint32_t)fy * (uint32_t)atlas_side + (uint32_t)fx;
      __hip_atomic_fetch_min(atlas + texel, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    rec = next;
    p = pn;
  }
}

template <int kThreads>
__global__ __launch_bounds__(kThreads) void ism_splat_kernel(
    const float4* __restrict__ points, const uint32_t* __restrict__ bin_offsets, uint32_t n_points,
    const float* __restrict__ vpl, int n_vpl, int n1d, int atlas_side, float z_far,
    uint32_t* __restrict__ atlas) {
  __shared__ Candidate cand[kMaxTest];
  splat_bin<kThreads>(cand, points, bin_offsets, n_points, vpl, n_vpl, n1d, atlas_side, z_far, atlas,
                      blockIdx.x);
}

// Batched splat: persistent workgroups over the (frame, bin) items.
__global__ __launch_bounds__(256) void batch_splat_kernel(const BatchFrame* __restrict__ frames, int n,
                                                          uint32_t total) {
  __shared__ Candidate cand[kMaxTest];
  for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
    const BatchFrame& f = frames[batch_find<&BatchFrame::first_bin>(frames, n, item)];
    splat_bin<256>(cand, reinterpret_cast<const float4*>(f.points), f.bin_offsets, f.n_points, f.g.vpl,
                   f.g.n_vpl, f.n1d, f.atlas_side, f.z_far, f.atlas, (int)(item - f.first_bin));
    __syncthreads();  // the next item restages the candidates
  }
}

// Batched clear: one 64 KiB chunk (4096 x 16 bytes) of one frame's atlas per item.
constexpr uint32_t kClearChunk = 4096;
__global__ __launch_bounds__(256) void batch_clear_kernel(const BatchFrame* __restrict__ frames, int n,
                                                          uint32_t total) {
  const uint4 ones = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
  for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
    const BatchFrame& f = frames[batch_find<&BatchFrame::first_clear>(frames, n, item)];
    const size_t n4 = (size_t)f.atlas_side * f.atlas_side / 4;
    const size_t c0 = (size_t)(item - f.first_clear) * kClearChunk;
    const size_t c1 = min(n4, c0 + kClearChunk);
    uint4* a = reinterpret_cast<uint4*>(f.atlas);
    for (size_t i = c0 + threadIdx.x; i < c1; i += 256) a[i] = ones;
  }
}

}  // namespace

hipError_t launch_atlas_clear(hipStream_t s, uint32_t* atlas, size_t texels) {
  const size_t n4 = texels / 4;  // atlas sides are multiples of 4
  size_t blocks = (n4 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(atlas_clear_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                     reinterpret_cast<uint4*>(atlas), n4);
  return hipGetLastError();
}

hipError_t launch_ism_splat(hipStream_t s, const float* points, const uint32_t* bin_offsets,
                            int n_bins, int n_points, const float* vpl_records, int n_vpl, int n1d,
                            int atlas_side, float z_far, uint32_t* atlas, int block) {
  if (n_bins <= 0) return hipSuccess;
  const float4* pts = reinterpret_cast<const float4*>(points);
  if (block == 64)
    hipLaunchKernelGGL(ism_splat_kernel<64>, dim3(n_bins), dim3(64), 0, s, pts, bin_offsets,
                       (uint32_t)n_points, vpl_records, n_vpl, n1d, atlas_side, z_far, atlas);
  else if (block == 128)
    hipLaunchKernelGGL(ism_splat_kernel<128>, dim3(n_bins), dim3(128), 0, s, pts, bin_offsets,
                       (uint32_t)n_points, vpl_records, n_vpl, n1d, atlas_side, z_far, atlas);
  else
    hipLaunchKernelGGL(ism_splat_kernel<256>, dim3(n_bins), dim3(256), 0, s, pts, bin_offsets,
                       (uint32_t)n_points, vpl_records, n_vpl, n1d, atlas_side, z_far, atlas);
  return hipGetLastError();
}

hipError_t launch_batch_clear(hipStream_t s, const BatchFrame* frames, int n, uint32_t total) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_clear_kernel, dim3(persistent_grid(total, 8)), dim3(256), 0, s, frames, n, total);
  return hipGetLastError();
}

hipError_t launch_batch_splat(hipStream_t s, const BatchFrame* frames, int n, uint32_t total) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_splat_kernel, dim3(persistent_grid(total, 8)), dim3(256), 0, s, frames, n, total);
  return hipGetLastError();
}

uint32_t batch_clear_chunks(int atlas_side) {
  const uint64_t n4 = (uint64_t)atlas_side * atlas_side / 4;
  return (uint32_t)((n4 + kClearChunk - 1) / kClearChunk);
}

}  // namespace mlgi
