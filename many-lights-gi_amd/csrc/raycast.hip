// Device G-buffer / RSM producer: the scene's BVH ray cast on the GPU, one
// lane per texel, one 8x8 tile per single-wave workgroup, with the host renderer's double-precision code
// (scene_trace.hpp), so both write the same texels. Replaces the
// reference's rasterisation of the G-buffer and RSM (RasterizationStage.cpp
// and model.frag) for inputs that change per request or per frame.
// Compiled with -ffp-contract=off like the host renderer.
//
// A 128^2 RSM is only 256 such tiles: single-wave workgroups spread them over
// every CU, where 16x16 workgroups would occupy a quarter of the chip. Each
// lane's traversal stack lives in LDS, entry i of lane l at i * 64 + l (one
// bank per lane), since a stack in registers indexed per lane costs a select
// chain per access.
#include "kernels.hpp"
#include "scene_trace.hpp"

namespace mlgi {
namespace {

using namespace mlgi_scene_impl;

struct LdsStack {
  StackEntry* lane;  // entry 0 of this lane
  __device__ void put(int i, const StackEntry& v) { lane[i * 64] = v; }
  __device__ StackEntry get(int i) const { return lane[i * 64]; }
};

__global__ __launch_bounds__(64) void raycast_kernel(SceneView s, RayCam rc, int w, int h, int rsm,
                                                      float* __restrict__ a, float* __restrict__ b,
                                                      float* __restrict__ c) {
  __shared__ StackEntry stack[kTraceStack * 64];
  const int x = blockIdx.x * 8 + (threadIdx.x & 7);
  const int y = blockIdx.y * 8 + (threadIdx.x >> 3);
  if (x >= w || y >= h) return;
  LdsStack st{stack + threadIdx.x};
  if (rsm == 1)
    shade_rsm(st, s, rc, x, y, w, h, a, b, c);  // diffuse, normal, depth
  else if (rsm == 2)
    shade_surface(st, s, rc, x, y, w, h, a, b);  // albedo, VSM moments
  else
    shade_gbuffer(st, s, rc, x, y, w, h, a, b);  // depth, normal
}

}  // namespace

hipError_t launch_raycast(hipStream_t st, const SceneView& s, const RayCam& rc, int w, int h, int rsm,
                          float* a, float* b, float* c) {
  const dim3 grid((w + 7) / 8, (h + 7) / 8);
  hipLaunchKernelGGL(raycast_kernel, grid, dim3(64), 0, st, s, rc, w, h, rsm, a, b, c);
  return hipGetLastError();
}

}  // namespace mlgi
