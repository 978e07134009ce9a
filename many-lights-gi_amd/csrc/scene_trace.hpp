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

// Binary BVH node of the build (scene.cpp); the ray caster walks the 4-wide
// form below.
struct Node {
  double lo[3], hi[3];
  // leaf: count > 0, left = first index; inner: left child (the smaller
  // centroids along the split axis), right = left + 1, count = -1 - axis
  int left, count;
};

// 4-wide BVH node: the binary tree with every inner node's grandchildren
// pulled up into its child slots, stored as structures of arrays (the four
// boxes and references of one node are one 128-byte line). Boxes are floats
// rounded outward, so each still contains its triangles.
struct alignas(16) Node4 {
  float lo[3][4], hi[3][4];
  int ref[4];    // inner child: node index; leaf child: first triangle (leaf order)
  int count[4];  // leaf child: triangles (1..4); inner child: 0; empty slot: -1
};

// Triangle geometry in leaf order for the intersection test: v0 and the two
// edges v1 - v0, v2 - v0 (computed once, in double, as the test would). The
// array carries three zero entries past the last triangle, so that a leaf's
// four slots can be loaded without a bounds test.
constexpr int kGeoPad = 3;
struct TriGeo {
  D3 v0, e1, e2;
};

struct Hit {
  double t;
  int tri;
};

// The scene as flat arrays (host vectors or device copies).
struct SceneView {
  const Node4* nodes4;
  const TriGeo* geo;  // leaf order
  const Tri* tris;    // scene order (normals, albedo, sampling)
  const int* order;   // leaf order -> scene order
  int n_nodes4, n_tris;
};

// Slab test of a float box (bounds taken exactly into double), without early
// exits: t0 only grows and t1 only shrinks, so testing once after the three
// axes gives the same answer.
MLGI_HD inline bool box_hit(const float lo_[3], const float hi_[3], D3 o, D3 inv, double tmax, double* tnear) {
  double t0 = 0.0, t1 = tmax;
  for (int k = 0; k < 3; ++k) {
    const double a = ((double)lo_[k] - comp(o, k)) * comp(inv, k);
    const double b = ((double)hi_[k] - comp(o, k)) * comp(inv, k);
    const double lo = (a > b) ? b : a, hi = (a > b) ? a : b;
    t0 = (t0 < lo) ? lo : t0;  // std::max(t0, lo)
    t1 = (hi < t1) ? hi : t1;  // std::min(t1, hi)
  }
  *tnear = t0;
  return !(t0 > t1);
}

// Child slot c of a 4-wide node.
MLGI_HD inline bool box4_hit(const Node4& nd, int c, D3 o, D3 inv, double tmax, double* tnear) {
  const float lo[3] = {nd.lo[0][c], nd.lo[1][c], nd.lo[2][c]};
  const float hi[3] = {nd.hi[0][c], nd.hi[1][c], nd.hi[2][c]};
  return box_hit(lo, hi, o, inv, tmax, tnear);
}

// Moller-Trumbore, two-sided
MLGI_HD inline bool tri_hit(const TriGeo& g, D3 o, D3 d, double tmax, double* th) {
  D3 p = cross(d, g.e2);
  double det = dot(g.e1, p);
  if (fabs(det) < 1e-14) return false;
  double inv = 1.0 / det;
  D3 s = o - g.v0;
  double u = dot(s, p) * inv;
  if (u < 0.0 || u > 1.0) return false;
  D3 q = cross(s, g.e1);
  double v = dot(d, q) * inv;
  if (v < 0.0 || u + v > 1.0) return false;
  double tt = dot(g.e2, q) * inv;
  if (tt <= 0.0 || tt >= tmax) return false;
  *th = tt;
  return true;
}

// Pending children of the traversal: at most three per level of the 4-wide
// tree, whose depth the build keeps within 16 (binary depth 32).
constexpr int kTraceStack = 48;

// A pending child: ref << 3 | count (inner node: count 0; leaf: 1..4
// triangles) and its box's entry distance rounded down to a float (the pop
// test then never drops a box the exact test would keep).
struct StackEntry {
  int rc;
  float tnear;
};

// Traversal stack in the caller's memory (host: a local array).
struct LocalStack {
  StackEntry e[kTraceStack];
  MLGI_HD void put(int i, const StackEntry& v) { e[i] = v; }
  MLGI_HD StackEntry get(int i) const { return e[i]; }
};

// Largest float <= x for x >= 0 (a rounded-up f is then > 0: one step down
// is the previous float).
MLGI_HD inline float round_down_f(double x) {
  float f = (float)x;
  if ((double)f > x) {
    union {
      float f;
      uint32_t u;
    } b;
    b.f = f;
    b.u -= 1u;
    f = b.f;
  }
  return f;
}

// Compare-exchange of the child order: by entry distance, then slot.
MLGI_HD inline void order2(double* t, int* slot, int* rc, int i, int j) {
  if (t[i] > t[j] || (t[i] == t[j] && slot[i] > slot[j])) {
    const double tt = t[i];
    t[i] = t[j];
    t[j] = tt;
    const int ss = slot[i];
    slot[i] = slot[j];
    slot[j] = ss;
    const int rr = rc[i];
    rc[i] = rc[j];
    rc[j] = rr;
  }
}

// Nearest hit with t < tmax. At an inner node the whole 128-byte node is read
// at once and its four child boxes tested, the hit children are ordered by
// entry distance (then slot; missed slots last), the nearest is visited and
// the others pushed, farthest first. A popped child is skipped if its box is
// entered beyond the current nearest hit. Leaves test their triangles in leaf
// order; among hits at exactly equal t the first one reached wins. All array
// indices are compile-time (registers, no private memory), and host and
// device run this same code, so they take the same decisions.
MLGI_HD inline D3 ray_inverse(D3 d) {
  return {1.0 / (d.x != 0 ? d.x : 1e-300), 1.0 / (d.y != 0 ? d.y : 1e-300), 1.0 / (d.z != 0 ? d.z : 1e-300)};
}

// Entry distance of a missed child slot: orders after every hit one.
constexpr double kMissed = 1e300;

// The ordering network of an inner step: hit children by entry distance,
// then slot; missed slots last.
MLGI_HD inline void order_children(double* tk, int* slot, int* rc) {
  order2(tk, slot, rc, 0, 1);
  order2(tk, slot, rc, 2, 3);
  order2(tk, slot, rc, 0, 2);
  order2(tk, slot, rc, 1, 3);
  order2(tk, slot, rc, 1, 2);
}

template <class Stack>
MLGI_HD inline Hit trace(const SceneView& s, D3 o, D3 d, double tmax, Stack& st) {
  Hit h{tmax, -1};
  const D3 inv = ray_inverse(d);
  int sp = 0;
  int ref = 0, count = 0;  // the root, an inner node
  while (true) {
    if (count == 0) {
      const Node4 nd = s.nodes4[ref];
      double tk[4];
      int slot[4], rc[4];
      for (int k = 0; k < 4; ++k) {
        double tn;
        const bool hit = box4_hit(nd, k, o, inv, h.t, &tn) && nd.count[k] >= 0;
        tk[k] = hit ? tn : kMissed;
        slot[k] = k;
        rc[k] = (nd.ref[k] << 3) | (nd.count[k] & 7);
      }
      order_children(tk, slot, rc);
      for (int i = 3; i >= 1; --i)  // never full (see kTraceStack); checked all the same
        if (tk[i] < kMissed && sp < kTraceStack) st.put(sp++, StackEntry{rc[i], round_down_f(tk[i])});
      if (tk[0] < kMissed) {
        ref = rc[0] >> 3;
        count = rc[0] & 7;
        continue;
      }
    } else {
      TriGeo g[4];
      for (int k = 0; k < 4; ++k) g[k] = s.geo[ref + k];  // padded: in bounds for every leaf
      for (int k = 0; k < 4; ++k) {
        double th;
        if (k < count && tri_hit(g[k], o, d, h.t, &th)) {
          h.t = th;
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

// One ray per thread over a traversal stack (the host renderer; the device's
// one-lane-per-texel form). The shading functions below call a tracer and
// store only where it says so (a cooperative device tracer has several lanes
// per ray, one of which stores).
template <class Stack>
struct StackTracer {
  Stack st;
  MLGI_HD Hit operator()(const SceneView& s, D3 o, D3 d, double tmax) { return trace(s, o, d, tmax, st); }
  MLGI_HD bool stores() const { return true; }
};

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
template <class Tracer>
MLGI_HD inline void shade_gbuffer(Tracer& tr, const SceneView& s, const RayCam& rc, int x, int y, int w, int h,
                                  float* depth, float* normal) {
  D3 o, d;
  pixel_ray(rc, x, y, w, h, &o, &d);
  Hit hit = tr(s, o, d, 1.0);
  if (!tr.stores()) return;
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
template <class Tracer>
MLGI_HD inline void shade_rsm(Tracer& tr, const SceneView& s, const RayCam& rc, int x, int y, int w, int h,
                              float* diffuse, float* normal, float* depth) {
  D3 o, d;
  pixel_ray(rc, x, y, w, h, &o, &d);
  Hit hit = tr(s, o, d, 1.0);
  if (!tr.stores()) return;
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
template <class Tracer>
MLGI_HD inline void shade_surface(Tracer& tr, const SceneView& s, const RayCam& rc, int x, int y, int w, int h, float* albedo,
                                  float* moments) {
  D3 o, d;
  pixel_ray(rc, x, y, w, h, &o, &d);
  Hit hit = tr(s, o, d, 1.0);
  if (!tr.stores()) return;
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
