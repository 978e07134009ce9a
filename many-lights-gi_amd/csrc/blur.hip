// Geometry-aware separable blur of the irradiance: gi/gi_blur.frag:24-72 as
// GIStage::blur runs it (GIStage.cpp:308-344), an x pass into a temporary and
// a y pass of the temporary, both weighted by the face normals and the
// linear depth (reprojection.glsl:6-13). Compiled with -ffp-contract=off and
// written in the oracle's operation order, so the result is bit-exact
// against its FP32 instantiation.
//
// Both passes are fused: a workgroup owns a 32 x 8 output tile, loads the
// irradiance, decoded normals and linear depths of the tile plus a 3-texel
// apron into LDS once, runs the x pass over the 14 rows the y pass needs
// (into LDS) and then the y pass. HBM traffic per pixel: the 28 input bytes
// (plus the apron, mostly served by L2) and the 12 output bytes, against two
// full-screen read/write passes in the reference. Out-of-range texels read 0
// (irradiance 0, normal (-1, -1, -1), depth 0), as the x pass's temporary
// does outside the image.
#include "kernels.hpp"

namespace mlgi {
namespace {

constexpr int kRadius = 3;
constexpr int kTw = 32, kTh = 8;
constexpr int kAw = kTw + 2 * kRadius, kAh = kTh + 2 * kRadius;  // apron region

__device__ __forceinline__ float linear_depth(float d, float A, float B) {
  const float zn = 2.0f * d - 1.0f;
  return -B / (A + zn);
}

__device__ __forceinline__ float max0(float x) { return 0.0f < x ? x : 0.0f; }  // std::max(0, x)

struct Center {
  float d, nx, ny, nz;
};

// One weighted sample of gi_blur.frag's processSample at distance i.
__device__ __forceinline__ void blur_sample(const Center& c, int i, float gr, float gg, float gb, float mx, float my,
                                            float mz, float ds, float& ar, float& ag, float& ab, float& fsum) {
  const float nf = 1.0f - max0(c.nx * mx + c.ny * my + c.nz * mz);
  const float dd = (ds - c.d) / 1.0f;
  const float nd = nf * 5.0f;
  const float factor = max0(((float)kRadius - (float)i + 1.0f) / 4.0f - dd * dd - nd * nd);
  fsum = fsum + factor;
  ar = ar + gr * factor;
  ag = ag + gg * factor;
  ab = ab + gb * factor;
}

__global__ __launch_bounds__(256) void gi_blur_kernel(const float* __restrict__ gi, const float* __restrict__ normal,
                                                      const float* __restrict__ depth, int w, int h, float A, float B,
                                                      float* __restrict__ out) {
  __shared__ float sg[3][kAh][kAw];   // irradiance
  __shared__ float sn[3][kAh][kAw];   // decoded normal
  __shared__ float sd[kAh][kAw];      // linear depth
  __shared__ float st[3][kAh][kTw];   // x-pass result
  const int tid = threadIdx.x;
  const int x0 = blockIdx.x * kTw, y0 = blockIdx.y * kTh;

  for (int i = tid; i < kAw * kAh; i += 256) {
    const int ly = i / kAw, lx = i % kAw;
    const int x = x0 + lx - kRadius, y = y0 + ly - kRadius;
    const bool in = x >= 0 && y >= 0 && x < w && y < h;
    const size_t q = in ? (size_t)y * w + x : 0;
    sg[0][ly][lx] = in ? gi[q * 3 + 0] : 0.0f;
    sg[1][ly][lx] = in ? gi[q * 3 + 1] : 0.0f;
    sg[2][ly][lx] = in ? gi[q * 3 + 2] : 0.0f;
    sn[0][ly][lx] = (in ? normal[q * 3 + 0] : 0.0f) * 2.0f - 1.0f;
    sn[1][ly][lx] = (in ? normal[q * 3 + 1] : 0.0f) * 2.0f - 1.0f;
    sn[2][ly][lx] = (in ? normal[q * 3 + 2] : 0.0f) * 2.0f - 1.0f;
    sd[ly][lx] = linear_depth(in ? depth[q] : 0.0f, A, B);
  }
  __syncthreads();

  // x pass over every apron row (rows outside the image stay 0)
  for (int i = tid; i < kTw * kAh; i += 256) {
    const int ly = i / kTw, tx = i % kTw;
    const int x = x0 + tx, y = y0 + ly - kRadius;
    const int cx = tx + kRadius;
    float ar = 0.0f, ag = 0.0f, ab = 0.0f;
    if (x < w && y >= 0 && y < h) {
      const Center c{sd[ly][cx], sn[0][ly][cx], sn[1][ly][cx], sn[2][ly][cx]};
      ar = sg[0][ly][cx];
      ag = sg[1][ly][cx];
      ab = sg[2][ly][cx];
      float fsum = 1.0f;
#pragma unroll
      for (int side = -1; side <= 1; side += 2)
#pragma unroll
        for (int k = 1; k <= kRadius; ++k) {
          const int sx = cx + side * k;
          blur_sample(c, k, sg[0][ly][sx], sg[1][ly][sx], sg[2][ly][sx], sn[0][ly][sx], sn[1][ly][sx],
                      sn[2][ly][sx], sd[ly][sx], ar, ag, ab, fsum);
        }
      ar = ar / fsum;
      ag = ag / fsum;
      ab = ab / fsum;
    }
    st[0][ly][tx] = ar;
    st[1][ly][tx] = ag;
    st[2][ly][tx] = ab;
  }
  __syncthreads();

  // y pass of the tile's pixels
  {
    const int tx = tid % kTw, ty = tid / kTw;
    const int x = x0 + tx, y = y0 + ty;
    if (x < w && y < h) {
      const int cy = ty + kRadius, cx = tx + kRadius;
      const Center c{sd[cy][cx], sn[0][cy][cx], sn[1][cy][cx], sn[2][cy][cx]};
      float ar = st[0][cy][tx], ag = st[1][cy][tx], ab = st[2][cy][tx];
      float fsum = 1.0f;
#pragma unroll
      for (int side = -1; side <= 1; side += 2)
#pragma unroll
        for (int k = 1; k <= kRadius; ++k) {
          const int sy = cy + side * k;
          blur_sample(c, k, st[0][sy][tx], st[1][sy][tx], st[2][sy][tx], sn[0][sy][cx], sn[1][sy][cx],
                      sn[2][sy][cx], sd[sy][cx], ar, ag, ab, fsum);
        }
      float* o = out + ((size_t)y * w + x) * 3;
      o[0] = ar / fsum;
      o[1] = ag / fsum;
      o[2] = ab / fsum;
    }
  }
}

}  // namespace

hipError_t launch_gi_blur(hipStream_t s, const float* gi, const float* normal, const float* depth, int w, int h,
                          const float* projection16, float* out) {
  const dim3 grid((w + kTw - 1) / kTw, (h + kTh - 1) / kTh);
  hipLaunchKernelGGL(gi_blur_kernel, grid, dim3(256), 0, s, gi, normal, depth, w, h, projection16[10],
                     projection16[14], out);
  return hipGetLastError();
}

}  // namespace mlgi
