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
  // leaf: count > 0, left = first index; inner: left child (the smaller
  // centroids along the split axis), right = left + 1, count = -1 - axis
  int left, count;
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

// The traversal keeps at most one pending node per tree level; the build
// keeps the BVH at most 32 levels deep (it falls back to median splits
// otherwise), and 32 entries are few enough for the device compiler to keep
// the stack in registers.
constexpr int kTraceStack = 32;

// Nearest hit with t < tmax. At an inner node both children's boxes are
// tested and the child on the ray's side of the split axis is visited first
// (the other pushed): the sign of the direction along the axis, not the box
// distances, decides, so the 64 coherent rays of a wave take the same order
// and do not diverge. A popped node's box is tested again against the
// current nearest hit. Among hits at exactly equal t the first one reached
// wins.
MLGI_HD inline Hit trace(const SceneView& s, D3 o, D3 d, double tmax) {
  Hit h{tmax, -1};
  D3 inv = {1.0 / (d.x != 0 ? d.x : 1e-300), 1.0 / (d.y != 0 ? d.y : 1e-300), 1.0 / (d.z != 0 ? d.z : 1e-300)};
  int stack[kTraceStack];
  int sp = 0;
  double tn;
  int node = box_hit(s.nodes[0], o, inv, h.t, &tn) ? 0 : -1;
  while (node >= 0) {
    const Node& nd = s.nodes[node];
    int next = -1;
    if (nd.count > 0) {
      for (int i = nd.left; i < nd.left + nd.count; ++i) {
        double th;
        const int ti = s.order[i];
        if (tri_hit(s.tris[ti], o, d, h.t, &th)) {
          h.t = th;
          h.tri = ti;
        }
      }
    } else {
      const bool left_first = comp(d, -1 - nd.count) >= 0.0;
      const int a = left_first ? nd.left : nd.left + 1, b = left_first ? nd.left + 1 : nd.left;
      double ta, tb;
      const bool ha = box_hit(s.nodes[a], o, inv, h.t, &ta);
      const bool hb = box_hit(s.nodes[b], o, inv, h.t, &tb);
      if (ha) {
        next = a;
        if (hb && sp < kTraceStack) stack[sp++] = b;
      } else if (hb) {
        next = b;
      }
    }
    while (next < 0 && sp > 0) {  // pop until a node still in front of the nearest hit
      const int c = stack[--sp];
      if (box_hit(s.nodes[c], o, inv, h.t, &tn)) next = c;
    }
    node = next;
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

// Deferred-shading inputs of texel (x, y) (model.frag:65-132 writes them in
// the reference): the surface albedo (0 for the background) and the VSM
// moments (d, d^2) of the hit's distance d from the camera's eye, which is
// the light's position when rendered from the light (model.frag:130 adds a
// screen-space derivative term to d^2 that a ray caster has no analogue
// of; it is left out). Background: moments (0, 0). Either output may be
// null.
MLGI_HD inline void shade_surface(const SceneView& s, const RayCam& rc, int x, int y, int w, int h, float* albedo,
                                  float* moments) {
  D3 o, d;
  pixel_ray(rc, x, y, w, h, &o, &d);
  Hit hit = trace(s, o, d, 1.0);
  const size_t p = (size_t)y * w + x;
  if (hit.tri < 0) {
    if (albedo) albedo[p * 3] = albedo[p * 3 + 1] = albedo[p * 3 + 2] = 0.0f;
    if (moments) moments[p * 2] = moments[p * 2 + 1] = 0.0f;
    return;
  }
  if (albedo)
    for (int k = 0; k < 3; ++k) albedo[p * 3 + k] = s.tris[hit.tri].albedo[k];
  if (moments) {
    const double dist = len((o + d * hit.t) - rc.eye);
    moments[p * 2] = (float)dist;
    moments[p * 2 + 1] = (float)(dist * dist);
  }
}

// ------------------------------------------------------------ point sampler
// Area-proportional surface points (the reference's tessellation + geometry
// pass output, ism.tesc:11-23, ism.geom:49-69), shared by the host sampler
// (scene.cpp) and the device sampler (points.hip) so both produce the same
// records. Randomness is counter based: draw k of point i is a splitmix64
// hash of (seed, i, k), so every point is independent of every other.

MLGI_HD inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
MLGI_HD inline double counter_uniform(uint64_t seed, uint64_t i, uint64_t k) {
  const uint64_t z = mix64(seed ^ mix64(i * 4 + k + 0x632BE59BD9B4E019ull));
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}
// random.glsl:34-66: random(vec3) = floatConstruct(hash(uvec3 bits))
MLGI_HD inline uint32_t jhash(uint32_t x) {
  x += (x << 10u);
  x ^= (x >> 6u);
  x += (x << 3u);
  x ^= (x >> 11u);
  x += (x << 15u);
  return x;
}
MLGI_HD inline float glsl_random3(float a, float b, float c) {
  uint32_t m = jhash(__builtin_bit_cast(uint32_t, a) ^ jhash(__builtin_bit_cast(uint32_t, b)) ^
                     jhash(__builtin_bit_cast(uint32_t, c)));
  m &= 0x007FFFFFu;
  m |= 0x3F800000u;
  return __builtin_bit_cast(float, m) - 1.0f;
}
// packUnorm4x8 component: round-to-nearest-even of clamp(c, 0, 1) * 255
MLGI_HD inline uint32_t unorm8(float c) {
  const float v = (c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c)) * 255.0f;
  return (uint32_t)rintf(v);
}
// std::upper_bound over cdf[0, n): first index with cdf[i] > u
MLGI_HD inline int upper_bound_d(const double* cdf, int n, double u) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cdf[mid] > u)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}
// Point i: a triangle drawn by area, a uniform point on it, its record
// (pos, packUnorm4x8(n * 0.5 + 0.5, r / 25)) and its bin
// int(random(pos) * n_bins), clamped to [0, n_bins).
MLGI_HD inline void sample_point(const Tri* tris, const double* cdf, int nt, double total, uint64_t seed,
                                 uint64_t i, float r_packed, int n_bins, float rec[4], int* bin) {
  const double u0 = counter_uniform(seed, i, 0) * total;
  int ti = upper_bound_d(cdf, nt, u0);
  if (ti >= nt) ti = nt - 1;
  const Tri& tr = tris[ti];
  const double r1 = sqrt(counter_uniform(seed, i, 1));
  const double r2 = counter_uniform(seed, i, 2);
  const D3 p = tr.v0 * (1.0 - r1) + tr.v1 * (r1 * (1.0 - r2)) + tr.v2 * (r1 * r2);
  const float px = (float)p.x, py = (float)p.y, pz = (float)p.z;
  const uint32_t bits = unorm8((float)tr.n.x * 0.5f + 0.5f) | (unorm8((float)tr.n.y * 0.5f + 0.5f) << 8) |
                        (unorm8((float)tr.n.z * 0.5f + 0.5f) << 16) | (unorm8(r_packed) << 24);
  rec[0] = px;
  rec[1] = py;
  rec[2] = pz;
  rec[3] = __builtin_bit_cast(float, bits);
  const int bi = (int)(glsl_random3(px, py, pz) * (float)n_bins);
  *bin = bi < 0 ? 0 : (bi > n_bins - 1 ? n_bins - 1 : bi);
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
// Running triangle areas (cdf[i] = sum of areas 0..i, in order); returns the total.
double mlgi_scene_area_cdf(const mlgi_scene* s, double* cdf);
// Packed point radius r / 25 for n points standing for ism_count points each.
float mlgi_scene_point_radius(double total_area, int64_t n_points, int32_t ism_count);
