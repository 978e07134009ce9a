// Device G-buffer / RSM producer: the scene's BVH ray cast on the GPU with
// the host renderer's double-precision code (scene_trace.hpp), so both write
// the same texels. Replaces the reference's rasterisation of the G-buffer and
// RSM (RasterizationStage.cpp and model.frag) for inputs that change per
// request or per frame. Compiled with -ffp-contract=off like the host
// renderer.
//
// Two forms, both single-wave workgroups (a small image then still spreads
// over every CU) with each ray's traversal stack in LDS (a stack in
// registers indexed per lane costs a select chain per access):
//   * one lane per texel, 8x8 texels per wave, for large images;
//   * four lanes per texel (a quad), 4x4 texels per wave, for images of at
//     most kQuadRays texels: a 128^2 RSM is 256 waves in the first form, one
//     per CU, with three of every CU's four SIMDs idle. The quad's lanes
//     test the four child boxes of an inner node, or the up to four
//     triangles of a leaf, one each, and exchange the results by DPP quad
//     broadcasts; every lane then runs the host traversal's ordering
//     network and leaf-order combination on the same values, so each lane
//     holds the host's traversal state and the texels are the host's.
#include "kernels.hpp"
#include "scene_trace.hpp"

namespace mlgi {
namespace {

using namespace mlgi_scene_impl;

constexpr size_t kQuadRays = size_t(1) << 18;

// Entry i of a ray's stack at base[i * stride].
struct LdsStack {
  StackEntry* base;
  int stride;
  __device__ void put(int i, const StackEntry& v) { base[i * stride] = v; }
  __device__ StackEntry get(int i) const { return base[i * stride]; }
};

// A quad's shared stack: lane 0 of the quad writes, all four read.
struct QuadStack {
  StackEntry* base;
  int stride;
  bool writer;
  __device__ void put(int i, const StackEntry& v) {
    if (writer) base[i * stride] = v;
  }
  __device__ StackEntry get(int i) const { return base[i * stride]; }
};

// Value of lane k of this lane's quad (DPP quad_perm [k, k, k, k]). Every lane
// of a quad is active wherever these are used: a quad's control flow is
// uniform.
template <int k>
__device__ __forceinline__ int quad_lane(int v) {
  return __builtin_amdgcn_update_dpp(0, v, k | (k << 2) | (k << 4) | (k << 6), 0xF, 0xF, false);
}
template <int k>
__device__ __forceinline__ double quad_lane(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)quad_lane<k>((int)(unsigned)u);
  const unsigned hi = (unsigned)quad_lane<k>((int)(unsigned)(u >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// trace() of scene_trace.hpp with the four box tests of an inner step and the
// four triangle tests of a leaf spread over the quad (lane q takes slot q).
// Leaf: lane q tests against the t at leaf entry; combining in slot order
// with `hit_k && t_k < t` is the host's sequential test against the running
// t (t_k itself does not depend on the bound).
__device__ Hit trace_quad(const SceneView& s, D3 o, D3 d, double tmax, QuadStack& st, int q) {
  Hit h{tmax, -1};
  const D3 inv = ray_inverse(d);
  int sp = 0;
  int ref = 0, count = 0;  // the root, an inner node
  while (true) {
    if (count == 0) {
      const float* nf = reinterpret_cast<const float*>(s.nodes4 + ref);
      const int* ni = reinterpret_cast<const int*>(s.nodes4 + ref);
      const float lo[3] = {nf[q], nf[4 + q], nf[8 + q]};
      const float hi[3] = {nf[12 + q], nf[16 + q], nf[20 + q]};
      const int cref = ni[24 + q], ccnt = ni[28 + q];
      double tn;
      const bool hit = box_hit(lo, hi, o, inv, h.t, &tn) && ccnt >= 0;
      const double key = hit ? tn : kMissed;
      const int rcq = (cref << 3) | (ccnt & 7);
      double tk[4] = {quad_lane<0>(key), quad_lane<1>(key), quad_lane<2>(key), quad_lane<3>(key)};
      int rc[4] = {quad_lane<0>(rcq), quad_lane<1>(rcq), quad_lane<2>(rcq), quad_lane<3>(rcq)};
      int slot[4] = {0, 1, 2, 3};
      order_children(tk, slot, rc);
      for (int i = 3; i >= 1; --i)  // never full (see kTraceStack); checked all the same
        if (tk[i] < kMissed && sp < kTraceStack) st.put(sp++, StackEntry{rc[i], round_down_f(tk[i])});
      if (tk[0] < kMissed) {
        ref = rc[0] >> 3;
        count = rc[0] & 7;
        continue;
      }
    } else {
      const TriGeo g = s.geo[ref + q];  // padded: in bounds for every leaf
      double th = 0.0;
      const int hit = (q < count && tri_hit(g, o, d, h.t, &th)) ? 1 : 0;
      const int hk[4] = {quad_lane<0>(hit), quad_lane<1>(hit), quad_lane<2>(hit), quad_lane<3>(hit)};
      const double tk[4] = {quad_lane<0>(th), quad_lane<1>(th), quad_lane<2>(th), quad_lane<3>(th)};
      for (int k = 0; k < 4; ++k) {
        if (hk[k] && tk[k] < h.t) {
          h.t = tk[k];
          h.tri = ref + k;
        }
      }
    }
    bool more = false;
    while (sp > 0) {
      const StackEntry e = st.get(--sp);
      if (!((double)e.tnear > h.t)) {
        ref = e.rc >> 3;
        count = e.rc & 7;
        more = true;
        break;
      }
    }
    if (!more) break;
  }
  if (h.tri >= 0) h.tri = s.order[h.tri];
  return h;
}

struct QuadTracer {
  QuadStack st;
  int q;
  __device__ Hit operator()(const SceneView& s, D3 o, D3 d, double tmax) { return trace_quad(s, o, d, tmax, st, q); }
  __device__ bool stores() const { return q == 0; }
};

template <class Tracer>
__device__ __forceinline__ void shade(Tracer& tr, const SceneView& s, const RayCam& rc, int x, int y, int w, int h,
                                      int mode, float* a, float* b, float* c) {
  if (mode == 1)
    shade_rsm(tr, s, rc, x, y, w, h, a, b, c);  // diffuse, normal, depth
  else if (mode == 2)
    shade_surface(tr, s, rc, x, y, w, h, a, b);  // albedo, VSM moments
  else
    shade_gbuffer(tr, s, rc, x, y, w, h, a, b);  // depth, normal
}

__global__ __launch_bounds__(64) void raycast_kernel(SceneView s, RayCam rc, int w, int h, int mode,
                                                      float* __restrict__ a, float* __restrict__ b,
                                                      float* __restrict__ c) {
  __shared__ StackEntry stack[kTraceStack * 64];
  const int x = blockIdx.x * 8 + (threadIdx.x & 7);
  const int y = blockIdx.y * 8 + (threadIdx.x >> 3);
  if (x >= w || y >= h) return;
  StackTracer<LdsStack> tr{LdsStack{stack + threadIdx.x, 64}};
  shade(tr, s, rc, x, y, w, h, mode, a, b, c);
}

__global__ __launch_bounds__(64) void raycast_quad_kernel(SceneView s, RayCam rc, int w, int h, int mode,
                                                           float* __restrict__ a, float* __restrict__ b,
                                                           float* __restrict__ c) {
  __shared__ StackEntry stack[kTraceStack * 16];
  const int q = threadIdx.x & 3, r = threadIdx.x >> 2;
  const int x = blockIdx.x * 4 + (r & 3);
  const int y = blockIdx.y * 4 + (r >> 2);
  if (x >= w || y >= h) return;  // the whole quad
  QuadTracer tr{QuadStack{stack + r, 16, q == 0}, q};
  shade(tr, s, rc, x, y, w, h, mode, a, b, c);
}

}  // namespace

hipError_t launch_raycast(hipStream_t st, const SceneView& s, const RayCam& rc, int w, int h, int rsm,
                          float* a, float* b, float* c) {
  if ((size_t)w * (size_t)h <= kQuadRays) {
    const dim3 grid((w + 3) / 4, (h + 3) / 4);
    hipLaunchKernelGGL(raycast_quad_kernel, grid, dim3(64), 0, st, s, rc, w, h, rsm, a, b, c);
  } else {
    const dim3 grid((w + 7) / 8, (h + 7) / 8);
    hipLaunchKernelGGL(raycast_kernel, grid, dim3(64), 0, st, s, rc, w, h, rsm, a, b, c);
  }
  return hipGetLastError();
}

}  // namespace mlgi
