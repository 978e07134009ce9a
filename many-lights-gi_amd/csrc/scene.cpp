// Host-side synthetic inputs (include/mlgi_scene.h): a seeded procedural
// scene, a BVH ray caster that writes the G-buffer and RSM in the formats the
// reference rasteriser produces (RasterizationStage.cpp:139-149,
// model.frag:93 faceNormal = N*0.5+0.5), and the area-proportional splat
// point set with the reference's point record and bin hash (ism.geom:49-68,
// random.glsl:34-66). Deterministic for a given seed and independent of the
// thread count.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/mlgi.h"
#include "../../include/mlgi_scene.h"
#include "scene_trace.hpp"

namespace {

using namespace mlgi_scene_impl;

// splitmix64 sequential generator (mix64 and the counter-based draws of the
// point sampler are in scene_trace.hpp)
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  double uniform() {
    s += 0x9E3779B97F4A7C15ull;
    return (double)(mix64(s) >> 11) * (1.0 / 9007199254740992.0);
  }
};
}  // namespace

struct mlgi_scene {
  std::vector<Tri> tris;
  std::vector<int> order;  // BVH leaf order
  std::vector<Node> nodes;    // binary BVH (build)
  std::vector<Node4> nodes4;  // 4-wide BVH (traversal)
  std::vector<TriGeo> geo;    // triangle geometry in leaf order
  double area = 0.0;
};

namespace {

const D3 kDefaultEye = {5.0, 5.0, 9.0};

// Depth of the subtree at node i (the traversal stack must hold it).
int bvh_depth(const mlgi_scene* s, int i) {
  const Node& nd = s->nodes[i];
  if (nd.count > 0) return 1;
  return 1 + std::max(bvh_depth(s, nd.left), bvh_depth(s, nd.left + 1));
}

void build_bvh_with(mlgi_scene* s, int sah_depth) {
  const int n = (int)s->tris.size();
  s->order.resize(n);
  for (int i = 0; i < n; ++i) s->order[i] = i;
  std::vector<D3> cen(n);
  for (int i = 0; i < n; ++i) cen[i] = (s->tris[i].v0 + s->tris[i].v1 + s->tris[i].v2) * (1.0 / 3.0);
  s->nodes.clear();
  s->nodes.reserve(2 * n + 2);
  struct Item {
    int node, begin, end, depth;
  };
  std::vector<Item> stack;
  s->nodes.push_back(Node{});
  // Root split by size: triangles spanning more than a quarter of the scene
  // (the room's walls) go to their own subtree, so the boxes of the median
  // splits below stay tight around the small geometry instead of every level
  // containing a wall (which each ray would then enter).
  double slo[3] = {1e300, 1e300, 1e300}, shi[3] = {-1e300, -1e300, -1e300};
  for (const Tri& t : s->tris)
    for (const D3* v : {&t.v0, &t.v1, &t.v2})
      for (int k = 0; k < 3; ++k) {
        slo[k] = std::min(slo[k], comp(*v, k));
        shi[k] = std::max(shi[k], comp(*v, k));
      }
  const double span = std::max(shi[0] - slo[0], std::max(shi[1] - slo[1], shi[2] - slo[2]));
  auto big = [&](int i) {
    const Tri& t = s->tris[i];
    for (int k = 0; k < 3; ++k) {
      const double lo = std::min(comp(t.v0, k), std::min(comp(t.v1, k), comp(t.v2, k)));
      const double hi = std::max(comp(t.v0, k), std::max(comp(t.v1, k), comp(t.v2, k)));
      if (hi - lo > 0.25 * span) return true;
    }
    return false;
  };
  const int n_big = (int)(std::stable_partition(s->order.begin(), s->order.end(), big) - s->order.begin());
  if (n_big > 0 && n_big < n && n > 4) {
    Node root{};
    for (int k = 0; k < 3; ++k) {
      root.lo[k] = slo[k];
      root.hi[k] = shi[k];
    }
    root.left = 1;
    root.count = -1;  // inner node; children: walls (1), the rest (2); axis 0 for the visiting order
    s->nodes[0] = root;
    s->nodes.push_back(Node{});
    s->nodes.push_back(Node{});
    stack.push_back({1, 0, n_big, 1});
    stack.push_back({2, n_big, n, 1});
  } else {
    stack.push_back({0, 0, n, 0});
  }
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    Node nd{};
    for (int k = 0; k < 3; ++k) {
      nd.lo[k] = 1e300;
      nd.hi[k] = -1e300;
    }
    double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
    for (int i = it.begin; i < it.end; ++i) {
      const Tri& t = s->tris[s->order[i]];
      for (const D3* v : {&t.v0, &t.v1, &t.v2})
        for (int k = 0; k < 3; ++k) {
          nd.lo[k] = std::min(nd.lo[k], comp(*v, k));
          nd.hi[k] = std::max(nd.hi[k], comp(*v, k));
        }
      for (int k = 0; k < 3; ++k) {
        clo[k] = std::min(clo[k], comp(cen[s->order[i]], k));
        chi[k] = std::max(chi[k], comp(cen[s->order[i]], k));
      }
    }
    const int cnt = it.end - it.begin;
    if (cnt <= 4) {
      nd.left = it.begin;
      nd.count = cnt;
      s->nodes[it.node] = nd;
      continue;
    }
    // Binned surface-area heuristic on the top levels: 16 centroid bins per
    // axis, the split plane of least (area x count) over both sides. Median
    // splits below (and where all centroids coincide or a side would be
    // empty), so the depth stays within sah_depth + log2(n / 4) + 1.
    constexpr int kBins = 16;
    int axis = -1, split = -1;
    double best = 1e300;
    for (int k = 0; k < 3 && it.depth < sah_depth; ++k) {
      const double ext = chi[k] - clo[k];
      if (!(ext > 0.0)) continue;
      int bc[kBins] = {};
      double blo[kBins][3], bhi[kBins][3];
      for (int b = 0; b < kBins; ++b)
        for (int q = 0; q < 3; ++q) {
          blo[b][q] = 1e300;
          bhi[b][q] = -1e300;
        }
      for (int i = it.begin; i < it.end; ++i) {
        const int ti = s->order[i];
        const int b = std::min(kBins - 1, (int)((comp(cen[ti], k) - clo[k]) / ext * kBins));
        bc[b]++;
        const Tri& t = s->tris[ti];
        for (const D3* v : {&t.v0, &t.v1, &t.v2})
          for (int q = 0; q < 3; ++q) {
            blo[b][q] = std::min(blo[b][q], comp(*v, q));
            bhi[b][q] = std::max(bhi[b][q], comp(*v, q));
          }
      }
      auto area = [](const double* lo, const double* hi) {
        if (lo[0] > hi[0]) return 0.0;
        const double x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
        return 2.0 * (x * y + y * z + z * x);
      };
      double rlo[kBins][3], rhi[kBins][3];
      int rc[kBins];
      double alo[3] = {1e300, 1e300, 1e300}, ahi[3] = {-1e300, -1e300, -1e300};
      int acc = 0;
      for (int b = kBins - 1; b >= 0; --b) {  // suffix boxes and counts
        for (int q = 0; q < 3; ++q) {
          alo[q] = std::min(alo[q], blo[b][q]);
          ahi[q] = std::max(ahi[q], bhi[b][q]);
          rlo[b][q] = alo[q];
          rhi[b][q] = ahi[q];
        }
        acc += bc[b];
        rc[b] = acc;
      }
      double llo[3] = {1e300, 1e300, 1e300}, lhi[3] = {-1e300, -1e300, -1e300};
      int lc = 0;
      for (int sp = 1; sp < kBins; ++sp) {  // left = bins [0, sp)
        for (int q = 0; q < 3; ++q) {
          llo[q] = std::min(llo[q], blo[sp - 1][q]);
          lhi[q] = std::max(lhi[q], bhi[sp - 1][q]);
        }
        lc += bc[sp - 1];
        if (lc == 0 || rc[sp] == 0) continue;
        const double cost = area(llo, lhi) * lc + area(rlo[sp], rhi[sp]) * rc[sp];
        if (cost < best) {
          best = cost;
          axis = k;
          split = sp;
        }
      }
    }
    int mid = -1;
    if (axis >= 0) {
      const double ext = chi[axis] - clo[axis];
      const int ax = axis, spl = split;
      const double c0 = clo[axis];
      mid = (int)(std::stable_partition(s->order.begin() + it.begin, s->order.begin() + it.end,
                                        [&](int ti) {
                                          return std::min(kBins - 1, (int)((comp(cen[ti], ax) - c0) / ext * kBins)) <
                                                 spl;
                                        }) -
                  s->order.begin());
    }
    if (mid <= it.begin || mid >= it.end) {
      axis = 0;
      for (int k = 1; k < 3; ++k)
        if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
      mid = (it.begin + it.end) / 2;
      std::nth_element(s->order.begin() + it.begin, s->order.begin() + mid, s->order.begin() + it.end,
                       [&](int a, int b) {
                         double ca = comp(cen[a], axis), cb = comp(cen[b], axis);
                         return ca < cb || (ca == cb && a < b);
                       });
    }
    nd.left = (int)s->nodes.size();
    nd.count = -1 - axis;  // inner node: split axis
    s->nodes[it.node] = nd;
    s->nodes.push_back(Node{});
    s->nodes.push_back(Node{});
    stack.push_back({nd.left, it.begin, mid, it.depth + 1});
    stack.push_back({nd.left + 1, mid, it.end, it.depth + 1});
  }
}

// Float bounds that contain the double ones.
float round_up_f(double x) {
  float f = (float)x;
  if ((double)f < x) f = std::nextafter(f, HUGE_VALF);
  return f;
}
float round_down_any(double x) {
  float f = (float)x;
  if ((double)f > x) f = std::nextafter(f, -HUGE_VALF);
  return f;
}

// 4-wide node of binary node b (an inner node, or a leaf root): its children,
// each inner child replaced by its own two children, in binary order.
int collapse(mlgi_scene* s, int b) {
  const int me = (int)s->nodes4.size();
  s->nodes4.push_back(Node4{});
  int slots[4], n = 0;
  const Node nd = s->nodes[b];
  if (nd.count > 0) {
    slots[n++] = b;
  } else {
    for (int c : {nd.left, nd.left + 1}) {
      const Node& cn = s->nodes[c];
      if (cn.count > 0) {
        slots[n++] = c;
      } else {
        slots[n++] = cn.left;
        slots[n++] = cn.left + 1;
      }
    }
  }
  Node4 q{};
  for (int k = 0; k < 4; ++k) {
    q.ref[k] = 0;
    q.count[k] = -1;
    for (int a = 0; a < 3; ++a) q.lo[a][k] = q.hi[a][k] = 0.0f;
  }
  for (int k = 0; k < n; ++k) {
    const Node c = s->nodes[slots[k]];
    for (int a = 0; a < 3; ++a) {
      q.lo[a][k] = round_down_any(c.lo[a]);
      q.hi[a][k] = round_up_f(c.hi[a]);
    }
    q.ref[k] = c.count > 0 ? c.left : collapse(s, slots[k]);
    q.count[k] = c.count > 0 ? c.count : 0;
  }
  s->nodes4[me] = q;
  return me;
}

// SAH on the top 8 levels; if the tree outgrows the traversal stack, which
// holds three pending children per 4-wide level and so needs a binary depth
// of at most 32 (only for far larger scenes than generated here), median
// splits only, which are ceil(log2(n / 4)) + 2 <= 32 deep. Then the 4-wide
// tree and the leaf-order triangle geometry.
void build_bvh(mlgi_scene* s) {
  build_bvh_with(s, 8);
  if (bvh_depth(s, 0) > 2 * kTraceStack / 3) build_bvh_with(s, 0);
  s->nodes4.clear();
  s->nodes4.reserve(s->nodes.size() / 2 + 2);
  collapse(s, 0);
  s->geo.assign(s->tris.size() + kGeoPad, TriGeo{});
  for (size_t i = 0; i < s->tris.size(); ++i) {
    const Tri& t = s->tris[s->order[i]];
    s->geo[i] = TriGeo{t.v0, t.v1 - t.v0, t.v2 - t.v0};
  }
}

// 4x4 column-major helpers in double
M4 mul(const M4& a, const M4& b) {
  M4 r{};
  for (int c = 0; c < 4; ++c)
    for (int rr = 0; rr < 4; ++rr) {
      double acc = 0;
      for (int k = 0; k < 4; ++k) acc += a.m[k * 4 + rr] * b.m[c * 4 + k];
      r.m[c * 4 + rr] = acc;
    }
  return r;
}
bool inverse(const M4& a, M4* out) {
  double m[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      m[r][c] = a.m[c * 4 + r];
      m[r][c + 4] = r == c ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(m[r][c]) > std::fabs(m[piv][c])) piv = r;
    if (std::fabs(m[piv][c]) < 1e-300) return false;
    for (int k = 0; k < 8; ++k) std::swap(m[c][k], m[piv][k]);
    double d = m[c][c];
    for (int k = 0; k < 8; ++k) m[c][k] /= d;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      double f = m[r][c];
      for (int k = 0; k < 8; ++k) m[r][k] -= f * m[c][k];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) out->m[c * 4 + r] = m[r][c + 4];
  return true;
}
void camera_matrices(const mlgi_camera* cam, M4* view, M4* proj) {
  D3 eye = {cam->eye[0], cam->eye[1], cam->eye[2]};
  D3 ctr = {cam->center[0], cam->center[1], cam->center[2]};
  D3 up = {cam->up[0], cam->up[1], cam->up[2]};
  D3 f = norm(ctr - eye);
  D3 s = norm(cross(f, up));
  D3 u = cross(s, f);
  M4 v{};
  v.m[0] = s.x; v.m[4] = s.y; v.m[8] = s.z;
  v.m[1] = u.x; v.m[5] = u.y; v.m[9] = u.z;
  v.m[2] = -f.x; v.m[6] = -f.y; v.m[10] = -f.z;
  v.m[12] = -dot(s, eye); v.m[13] = -dot(u, eye); v.m[14] = dot(f, eye);
  v.m[15] = 1.0;
  *view = v;
  const double th = std::tan((double)cam->fovy_deg * M_PI / 180.0 / 2.0);
  const double n = cam->z_near, fa = cam->z_far;
  M4 p{};
  p.m[0] = 1.0 / ((double)cam->aspect * th);
  p.m[5] = 1.0 / th;
  p.m[10] = -(fa + n) / (fa - n);
  p.m[11] = -1.0;
  p.m[14] = -(2.0 * fa * n) / (fa - n);
  *proj = p;
}

void to_float(const M4& a, float* o) {
  for (int i = 0; i < 16; ++i) o[i] = (float)a.m[i];
}

// Worker threads of the host renderer and sampler: `t` when positive, else
// OMP_NUM_THREADS when set (the process's share of the host), else the
// hardware's thread count (at most 64).
int nthreads(int t) {
  if (t > 0) return t;
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int n = std::atoi(e);
    if (n > 0) return std::min(n, 64);
  }
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)std::min(h, 64u) : 4;
}

template <class F>
void parallel_rows(int rows, int threads, F f) {
  const int nt = std::max(1, std::min(threads, rows));
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([=]() {
      for (int y = t; y < rows; y += nt) f(y);
    });
  for (auto& th : pool) th.join();
}

RayCam ray_cam(const mlgi_camera* cam) {
  M4 v, p;
  camera_matrices(cam, &v, &p);
  RayCam rc;
  rc.vp = mul(p, v);
  inverse(rc.vp, &rc.inv_vp);
  rc.eye = {cam->eye[0], cam->eye[1], cam->eye[2]};
  rc.axis = norm(D3{cam->center[0], cam->center[1], cam->center[2]} - rc.eye);
  rc.cos_half = std::cos((double)cam->fovy_deg * M_PI / 180.0 / 2.0);
  return rc;
}

SceneView view_of(const mlgi_scene* s);

}  // namespace

extern "C" {

int mlgi_scene_create(uint64_t seed, int32_t n_random, mlgi_scene** out) {
  if (!out || n_random < 0) return MLGI_ERR_INVALID_ARGUMENT;
  auto* s = new mlgi_scene();
  Rng rng(seed);
  const double L = 10.0;
  auto quad = [&](D3 a, D3 b, D3 c, D3 d, D3 n) {
    for (int k = 0; k < 2; ++k) {
      Tri t;
      t.v0 = a;
      t.v1 = k == 0 ? b : c;
      t.v2 = k == 0 ? c : d;
      t.n = n;
      for (int i = 0; i < 3; ++i) t.albedo[i] = (float)(0.2 + 0.6 * rng.uniform());
      s->tris.push_back(t);
    }
  };
  quad({0, 0, 0}, {L, 0, 0}, {L, 0, L}, {0, 0, L}, {0, 1, 0});    // floor
  quad({0, L, 0}, {L, L, 0}, {L, L, L}, {0, L, L}, {0, -1, 0});   // ceiling
  quad({0, 0, 0}, {0, L, 0}, {0, L, L}, {0, 0, L}, {1, 0, 0});    // x = 0
  quad({L, 0, 0}, {L, L, 0}, {L, L, L}, {L, 0, L}, {-1, 0, 0});   // x = L
  quad({0, 0, 0}, {L, 0, 0}, {L, L, 0}, {0, L, 0}, {0, 0, 1});    // z = 0
  quad({0, 0, L}, {L, 0, L}, {L, L, L}, {0, L, L}, {0, 0, -1});   // z = L
  auto dir = [&]() {
    double z = 2.0 * rng.uniform() - 1.0, ph = 2.0 * M_PI * rng.uniform();
    double r = std::sqrt(std::max(0.0, 1.0 - z * z));
    return D3{r * std::cos(ph), r * std::sin(ph), z};
  };
  int made = 0;
  while (made < n_random) {
    D3 v0 = {0.5 + 9.0 * rng.uniform(), 0.5 + 9.0 * rng.uniform(), 0.5 + 9.0 * rng.uniform()};
    D3 e1 = dir() * (0.1 + 0.4 * rng.uniform());
    D3 e2 = dir() * (0.1 + 0.4 * rng.uniform());
    float alb[3];
    for (int i = 0; i < 3; ++i) alb[i] = (float)(0.2 + 0.6 * rng.uniform());
    D3 c = cross(e1, e2);
    if (len(c) < 1e-4) continue;
    if (len(v0 - kDefaultEye) < 1.0) continue;  // keep the default camera clear
    Tri t;
    t.v0 = v0;
    t.v1 = v0 + e1;
    t.v2 = v0 + e2;
    t.n = norm(c);
    std::memcpy(t.albedo, alb, sizeof alb);
    s->tris.push_back(t);
    ++made;
  }
  for (const Tri& t : s->tris) s->area += 0.5 * len(cross(t.v1 - t.v0, t.v2 - t.v0));
  build_bvh(s);
  *out = s;
  return MLGI_OK;
}

void mlgi_scene_destroy(mlgi_scene* s) { delete s; }

int32_t mlgi_scene_triangle_count(const mlgi_scene* s) { return s ? (int32_t)s->tris.size() : 0; }

double mlgi_scene_total_area(const mlgi_scene* s) { return s ? s->area : 0.0; }

int mlgi_scene_triangles(const mlgi_scene* s, float* v, float* n, float* a) {
  if (!s) return MLGI_ERR_INVALID_ARGUMENT;
  for (size_t i = 0; i < s->tris.size(); ++i) {
    const Tri& t = s->tris[i];
    if (v) {
      const D3* p[3] = {&t.v0, &t.v1, &t.v2};
      for (int k = 0; k < 3; ++k) {
        v[i * 9 + k * 3 + 0] = (float)p[k]->x;
        v[i * 9 + k * 3 + 1] = (float)p[k]->y;
        v[i * 9 + k * 3 + 2] = (float)p[k]->z;
      }
    }
    if (n) {
      n[i * 3 + 0] = (float)t.n.x;
      n[i * 3 + 1] = (float)t.n.y;
      n[i * 3 + 2] = (float)t.n.z;
    }
    if (a) std::memcpy(a + i * 3, t.albedo, sizeof t.albedo);
  }
  return MLGI_OK;
}

void mlgi_scene_default_camera(float aspect, mlgi_camera* c) {
  if (!c) return;
  const float eye[3] = {(float)kDefaultEye.x, (float)kDefaultEye.y, (float)kDefaultEye.z};
  const float ctr[3] = {5.0f, 4.0f, 0.0f};
  const float up[3] = {0.0f, 1.0f, 0.0f};
  std::memcpy(c->eye, eye, sizeof eye);
  std::memcpy(c->center, ctr, sizeof ctr);
  std::memcpy(c->up, up, sizeof up);
  c->fovy_deg = 60.0f;
  c->aspect = aspect;
  c->z_near = 0.05f;
  c->z_far = 50.0f;
}

void mlgi_scene_light_pose(uint64_t seed, mlgi_camera* c) {
  if (!c) return;
  Rng rng(seed ^ 0x5107A11C0FFEEull);
  c->eye[0] = (float)(2.0 + 6.0 * rng.uniform());
  c->eye[1] = (float)(8.5 + 1.0 * rng.uniform());
  c->eye[2] = (float)(2.0 + 6.0 * rng.uniform());
  c->center[0] = (float)(1.0 + 8.0 * rng.uniform());
  c->center[1] = (float)(0.0 + 2.0 * rng.uniform());
  c->center[2] = (float)(1.0 + 8.0 * rng.uniform());
  c->up[0] = 0.0f;
  c->up[1] = 0.0f;
  c->up[2] = 1.0f;
  c->fovy_deg = 45.0f;
  c->aspect = 1.0f;
  c->z_near = 0.05f;
  c->z_far = 50.0f;
}

int mlgi_camera_matrices(const mlgi_camera* cam, float* view16, float* proj16, float* vp_inv16) {
  if (!cam) return MLGI_ERR_INVALID_ARGUMENT;
  M4 v, p;
  camera_matrices(cam, &v, &p);
  if (view16) to_float(v, view16);
  if (proj16) to_float(p, proj16);
  if (vp_inv16) {
    M4 inv;
    if (!inverse(mul(p, v), &inv)) return MLGI_ERR_INVALID_ARGUMENT;
    to_float(inv, vp_inv16);
  }
  return MLGI_OK;
}

int mlgi_scene_render_gbuffer(const mlgi_scene* s, const mlgi_camera* cam, int32_t w, int32_t h,
                              int32_t threads, float* depth, float* normal) {
  if (!s || !cam || w <= 0 || h <= 0 || !depth || !normal) return MLGI_ERR_INVALID_ARGUMENT;
  const RayCam rc = ray_cam(cam);
  const SceneView sv = view_of(s);
  parallel_rows(h, nthreads(threads), [&](int y) {
    StackTracer<LocalStack> tr;
    for (int x = 0; x < w; ++x) shade_gbuffer(tr, sv, rc, x, y, w, h, depth, normal);
  });
  return MLGI_OK;
}

int mlgi_scene_render_rsm(const mlgi_scene* s, const mlgi_camera* light, int32_t w, int32_t h,
                          int32_t threads, float* diffuse, float* normal, float* depth,
                          float* inv_bvp16) {
  if (!s || !light || w <= 0 || h <= 0 || !diffuse || !normal || !depth)
    return MLGI_ERR_INVALID_ARGUMENT;
  const RayCam rc = ray_cam(light);
  const SceneView sv = view_of(s);
  parallel_rows(h, nthreads(threads), [&](int y) {
    StackTracer<LocalStack> tr;
    for (int x = 0; x < w; ++x) shade_rsm(tr, sv, rc, x, y, w, h, diffuse, normal, depth);
  });
  if (inv_bvp16 && !mlgi_scene_inv_biased_vp(light, inv_bvp16)) return MLGI_ERR_INVALID_ARGUMENT;
  return MLGI_OK;
}

int mlgi_scene_render_surface(const mlgi_scene* s, const mlgi_camera* cam, int32_t w, int32_t h, int32_t threads,
                              float* albedo, float* moments) {
  if (!s || !cam || w <= 0 || h <= 0 || (!albedo && !moments)) return MLGI_ERR_INVALID_ARGUMENT;
  const RayCam rc = ray_cam(cam);
  const SceneView sv = view_of(s);
  parallel_rows(h, nthreads(threads), [&](int y) {
    StackTracer<LocalStack> tr;
    for (int x = 0; x < w; ++x) shade_surface(tr, sv, rc, x, y, w, h, albedo, moments);
  });
  return MLGI_OK;
}

}  // extern "C"

double mlgi_scene_area_cdf(const mlgi_scene* s, double* cdf) {
  double acc = 0.0;
  for (size_t i = 0; i < s->tris.size(); ++i) {
    const Tri& t = s->tris[i];
    acc += 0.5 * len(cross(t.v1 - t.v0, t.v2 - t.v0));
    cdf[i] = acc;
  }
  return acc;
}

float mlgi_scene_point_radius(double total_area, int64_t n_points, int32_t ism_count) {
  // each point stands for ism_count points: radius boosted by sqrt(ism_count) (ism.geom:60-63)
  const double r_world = std::sqrt(total_area / ((double)std::max<int64_t>(n_points, 1) * M_PI)) *
                         std::sqrt((double)ism_count);
  return (float)(r_world / 25.0);
}

extern "C" int mlgi_scene_sample_points(const mlgi_scene* s, int64_t n_points, int32_t n_bins,
                                        int32_t ism_count, uint64_t seed, int32_t threads,
                                        float* records, uint32_t* bin_offsets) {
  if (!s || n_points < 0 || n_points > (1ll << 31) - 1 || n_bins <= 0 || ism_count <= 0 ||
      (!records && n_points) || !bin_offsets)
    return MLGI_ERR_INVALID_ARGUMENT;
  const size_t nt = s->tris.size();
  std::vector<double> cdf(nt);
  const double total = mlgi_scene_area_cdf(s, cdf.data());
  const float r_packed = mlgi_scene_point_radius(total, n_points, ism_count);
  std::vector<float> tmp((size_t)n_points * 4);
  std::vector<int32_t> bin((size_t)n_points);
  const int nth = nthreads(threads);
  const int64_t chunk = (n_points + nth - 1) / std::max(nth, 1);
  std::vector<std::thread> pool;
  for (int t = 0; t < nth; ++t) {
    pool.emplace_back([&, t]() {
      const int64_t b = t * chunk, e = std::min<int64_t>(n_points, b + chunk);
      for (int64_t i = b; i < e; ++i)
        sample_point(s->tris.data(), cdf.data(), (int)nt, total, seed, (uint64_t)i, r_packed, n_bins, &tmp[i * 4],
                     &bin[i]);
    });
  }
  for (auto& th : pool) th.join();
  std::vector<uint32_t> cnt(n_bins + 1, 0);
  for (int64_t i = 0; i < n_points; ++i) cnt[bin[i] + 1]++;
  for (int b = 0; b < n_bins; ++b) cnt[b + 1] += cnt[b];
  std::memcpy(bin_offsets, cnt.data(), sizeof(uint32_t) * (n_bins + 1));
  std::vector<uint32_t> pos(cnt.begin(), cnt.end() - 1);
  for (int64_t i = 0; i < n_points; ++i) {
    uint32_t dst = pos[bin[i]]++;
    std::memcpy(records + (size_t)dst * 4, &tmp[i * 4], 16);
  }
  return MLGI_OK;
}

namespace {
SceneView view_of(const mlgi_scene* s) {
  return SceneView{s->nodes4.data(), s->geo.data(), s->tris.data(), s->order.data(), (int)s->nodes4.size(),
                   (int)s->tris.size()};
}
}  // namespace

SceneView mlgi_scene_view(const mlgi_scene* s) { return view_of(s); }

bool mlgi_scene_ray_cam(const mlgi_camera* cam, RayCam* out) {
  if (!cam || !out) return false;
  *out = ray_cam(cam);
  return true;
}

bool mlgi_scene_inv_biased_vp(const mlgi_camera* light, float* out16) {
  const RayCam rc = ray_cam(light);
  M4 bias{};
  bias.m[0] = 0.5; bias.m[5] = 0.5; bias.m[10] = 0.5;
  bias.m[12] = 0.5; bias.m[13] = 0.5; bias.m[14] = 0.5; bias.m[15] = 1.0;
  M4 inv;
  if (!inverse(mul(bias, rc.vp), &inv)) return false;
  to_float(inv, out16);
  return true;
}
