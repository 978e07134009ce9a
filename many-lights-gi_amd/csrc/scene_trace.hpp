// Ray casting of the procedural scene, shared by the host renderer
// (scene.cpp) and the device renderer (raycast.hip): the same double
// precision code (both compiled with -ffp-contract=off) so that both write
// the same G-buffer and RSM texels.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define MLGI_HD __host__ __device__
#else
#define MLGI_HD
#endif

namespace mlgi_scene_impl {

struct D3 {
  double x, y, z;
};
MLGI_HD inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
MLGI_HD inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
MLGI_HD inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
MLGI_HD inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MLGI_HD inline D3 cross(D3 a, D3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
MLGI_HD inline double len(D3 a) { return sqrt(dot(a, a)); }
MLGI_HD inline D3 norm(D3 a) { return a * (1.0 / len(a)); }
MLGI_HD inline double comp(D3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

struct Tri {
  D3 v0, v1, v2, n;
  float albedo[3];
};

struct Node {
  double lo[3], hi[3];
  int left, count;  // leaf: count > 0, left = first index; inner: left child, right = left+1
};

struct Hit {
  double t;
  int tri;
};

// The BVH as flat arrays (host vectors or device copies).
struct SceneView {
  const Node* nodes;
  const Tri* tris;
  const int* order;  // leaf triangle order
  int n_nodes, n_tris;
};

MLGI_HD inline bool box_hit(const Node& nd, D3 o, D3 inv, double tmax, double* tnear) {
  double t0 = 0.0, t1 = tmax;
  for (int k = 0; k < 3; ++k) {
    double a = (nd.lo[k] - comp(o, k)) * comp(inv, k);
    double b = (nd.hi[k] - comp(o, k)) * comp(inv, k);
    if (a > b) {
      const double t = a;
      a = b;
      b = t;
    }
    t0 = (t0 < a) ? a : t0;  // std::max(t0, a)
    t1 = (b < t1) ? b : t1;  // std::min(t1, b)
    if (t0 > t1) return false;
  }
  *tnear = t0;
  return true;
}

// Moller-Trumbore, two-sided
MLGI_HD inline bool tri_hit(const Tri& t, D3 o, D3 d, double tmax, double* th) {
  D3 e1 = t.v1 - t.v0, e2 = t.v2 - t.v0;
  D3 p = cross(d, e2);
  double det = dot(e1, p);
  if (fabs(det) < 1e-14) return false;
  double inv = 1.0 / det;
  D3 s = o - t.v0;
  double u = dot(s, p) * inv;
  if (u < 0.0 || u > 1.0) return false;
  D3 q = cross(s, e1);
  double v = dot(d, q) * inv;
  if (v < 0.0 || u + v > 1.0) return false;
  double tt = dot(e2, q) * inv;
  if (tt <= 0.0 || tt >= tmax) return false;
  *th = tt;
  return true;
}

constexpr int kTraceStack = 64;  // the median-split BVH of 2^30 triangles is < 30 deep

// Nearest hit (smallest t, then smallest triangle index) with t < tmax.
MLGI_HD inline Hit trace(const SceneView& s, D3 o, D3 d, double tmax) {
  Hit h{tmax, -1};
  D3 inv = {1.0 / (d.x != 0 ? d.x : 1e-300), 1.0 / (d.y != 0 ? d.y : 1e-300), 1.0 / (d.z != 0 ? d.z : 1e-300)};
  int stack[kTraceStack];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const Node& nd = s.nodes[stack[--sp]];
    double tn;
    if (!box_hit(nd, o, inv, h.t, &tn)) continue;
    if (nd.count > 0) {
      for (int i = nd.left; i < nd.left + nd.count; ++i) {
        double th;
        int ti = s.order[i];
        if (tri_hit(s.tris[ti], o, d, h.t, &th) && (th < h.t || (th == h.t && ti < h.tri))) {
          h.t = th;
          h.tri = ti;
        }
      }
    } else if (sp + 2 <= kTraceStack) {
      stack[sp++] = nd.left;
      stack[sp++] = nd.left + 1;
    }
  }
  return h;
}

// 4x4 column-major, double
struct M4 {
  double m[16];
};
MLGI_HD inline void xform(const M4& a, const double v[4], double o[4]) {
  for (int r = 0; r < 4; ++r) {
    double acc = 0;
    for (int c = 0; c < 4; ++c) acc += a.m[c * 4 + r] * v[c];
    o[r] = acc;
  }
}

struct RayCam {
  M4 vp, inv_vp;
  D3 eye, axis;
  double cos_half;  // spot cone (RSM only)
};

// Ray through the pixel centre from the near plane to the far plane.
MLGI_HD inline void pixel_ray(const RayCam& rc, int x, int y, int w, int h, D3* o, D3* d) {
  double nx = ((x + 0.5) / w) * 2.0 - 1.0, ny = ((y + 0.5) / h) * 2.0 - 1.0;
  double a[4] = {nx, ny, -1.0, 1.0}, b[4] = {nx, ny, 1.0, 1.0}, pa[4], pb[4];
  xform(rc.inv_vp, a, pa);
  xform(rc.inv_vp, b, pb);
  D3 p0 = {pa[0] / pa[3], pa[1] / pa[3], pa[2] / pa[3]};
  D3 p1 = {pb[0] / pb[3], pb[1] / pb[3], pb[2] / pb[3]};
  *o = p0;
  *d = p1 - p0;
}

MLGI_HD inline double window_depth(const RayCam& rc, D3 p) {
  double v[4] = {p.x, p.y, p.z, 1.0}, c[4];
  xform(rc.vp, v, c);
  return 0.5 * (c[2] / c[3]) + 0.5;
}

// G-buffer texel (x, y): window depth and face normal * 0.5 + 0.5, the
// normal turned towards the viewer; background = depth 1, normal 0.5.
MLGI_HD inline void shade_gbuffer(const SceneView& s, const RayCam& rc, int x, int y, int w, int h,
                                  float* depth, float* normal) {
  D3 o, d;
  pixel_ray(rc, x, y, w, h, &o, &d);
  Hit hit = trace(s, o, d, 1.0);
  const size_t p = (size_t)y * w + x;
  if (hit.tri < 0) {
    depth[p] = 1.0f;
    normal[p * 3] = normal[p * 3 + 1] = normal[p * 3 + 2] = 0.5f;
    return;
  }
  D3 n = s.tris[hit.tri].n;
  if (dot(n, d) > 0) n = n * -1.0;
  depth[p] = (float)window_depth(rc, o + d * hit.t);
  normal[p * 3 + 0] = (float)n.x * 0.5f + 0.5f;
  normal[p * 3 + 1] = (float)n.y * 0.5f + 0.5f;
  normal[p * 3 + 2] = (float)n.z * 0.5f + 0.5f;
}

// RSM texel (x, y) of a spot light: as the G-buffer, plus the albedo inside
// the cone (0 outside) as the diffuse flux channel.
MLGI_HD inline void shade_rsm(const SceneView& s, const RayCam& rc, int x, int y, int w, int h,
                              float* diffuse, float* normal, float* depth) {
  D3 o, d;
  pixel_ray(rc, x, y, w, h, &o, &d);
  Hit hit = trace(s, o, d, 1.0);
  const size_t p = (size_t)y * w + x;
  if (hit.tri < 0) {
    depth[p] = 1.0f;
    for (int k = 0; k < 3; ++k) {
      normal[p * 3 + k] = 0.5f;
      diffuse[p * 3 + k] = 0.0f;
    }
    return;
  }
  const Tri& t = s.tris[hit.tri];
  D3 n = t.n;
  if (dot(n, d) > 0) n = n * -1.0;
  const bool in_cone = dot(norm(d), rc.axis) >= rc.cos_half;
  depth[p] = (float)window_depth(rc, o + d * hit.t);
  normal[p * 3 + 0] = (float)n.x * 0.5f + 0.5f;
  normal[p * 3 + 1] = (float)n.y * 0.5f + 0.5f;
  normal[p * 3 + 2] = (float)n.z * 0.5f + 0.5f;
  for (int k = 0; k < 3; ++k) diffuse[p * 3 + k] = in_cone ? t.albedo[k] : 0.0f;
}

}  // namespace mlgi_scene_impl

struct mlgi_scene;
// Host-side access for the device uploader (scene.cpp).
mlgi_scene_impl::SceneView mlgi_scene_view(const mlgi_scene* s);
// Ray camera of a camera / light: view-projection and inverse (double),
// eye, view axis, cos of the half cone angle.
bool mlgi_scene_ray_cam(const struct mlgi_camera* cam, mlgi_scene_impl::RayCam* out);
// inverse(bias * P * V) of a light as floats (column-major).
bool mlgi_scene_inv_biased_vp(const struct mlgi_camera* light, float* out16);
