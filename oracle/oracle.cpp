// Oracle for the many-lights GI hot path -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference GLSL, written from the shader text.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
// it, as the checker; the product library never links or calls it.
//
// Each stage is one template instantiated twice:
//   T = double : FP64 transcription (accuracy oracle)
//   T = float  : FP32 twin, fixed evaluation order, compiled with
//                -ffp-contract=off (bit-exact oracle for integer outputs)
//
// Choices for behaviour the GLSL leaves to the driver (stated once here):
//   * normalize(x) = x / sqrt(dot(x,x)); length(x) = sqrt(dot(x,x));
//     dot accumulates x, y, z left to right; mat*vec accumulates columns
//     in order (same as dot of the rows).
//   * pow(2, -level) is the exact power of two.
//   * float->uint conversion truncates and saturates at 0 (GCN cvt_u32).
//   * packHalf2x16 rounds to nearest even; packUnorm4x8 rounds to nearest
//     even; the R16 unorm store is rint(clamp(x,0,1) * 65535).
//   * out-of-range texelFetch / imageLoad return 0, out-of-range
//     imageAtomicMin is a no-op; NEAREST sampling wraps (GL_REPEAT).
//   * the VPL matrix product uses fused multiply-adds in a fixed order,
//     depth column first (fmaf in the FP32 twin), so the unprojection of a
//     perspective RSM stays well conditioned.
//   * constants written as float literals in GLSL (0.01, 0.6, 3.14, 1.3,
//     0.02) keep their binary32 value in the FP64 transcription.

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>
#include <atomic>

namespace {

template <class T> struct V3 {
  T x, y, z;
};
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> sub(V3<T> a, V3<T> b) { return v3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return v3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) {
  T l = length(a);
  return v3<T>(a.x / l, a.y / l, a.z / l);
}

inline float madd(float a, float b, float c) { return std::fma(a, b, c); }
inline double madd(double a, double b, double c) { return std::fma(a, b, c); }

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// GLSL uint(float): truncation; negative values saturate to 0 as on GCN.
template <class T> inline uint32_t to_uint(T v) {
  if (!(v > T(0))) return 0u;
  if (v >= T(4294967296.0)) return 0xFFFFFFFFu;
  return (uint32_t)v;
}

// Runs f(row) for every row in [0, n) on `threads` threads (rows taken from
// a shared counter). Rows must be independent.
template <class F> void parallel_rows(int n, int threads, F f) {
  std::atomic<int> next{0};
  auto worker = [&]() {
    for (int r = next.fetch_add(1); r < n; r = next.fetch_add(1)) f(r);
  };
  const int nt = std::max(1, std::min(threads, n));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
}

// dst = min(dst, v), atomically when other threads update the same array.
inline void store_min(uint32_t* dst, uint32_t v, bool atomic) {
  if (!atomic) {
    if (v < *dst) *dst = v;
    return;
  }
  uint32_t cur = __atomic_load_n(dst, __ATOMIC_RELAXED);
  while (v < cur && !__atomic_compare_exchange_n(dst, &cur, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
}

// ---------------------------------------------------------------- packing
// floatpacking.glsl:26-38 pack3UNToFloat, kept as uint32 bits.
template <class T> uint32_t pack3un(T x, T y, T z) {
  uint32_t u = to_uint(x * T(1023.0) + T(0.5));
  u |= to_uint(y * T(1023.0) + T(0.5)) << 10;
  u |= to_uint(z * T(1021.0) + T(1.5)) << 20;
  return u;
}
// floatpacking.glsl:40-44
template <class T> uint32_t pack3sn(T x, T y, T z) {
  return pack3un<T>(x * T(0.5) + T(0.5), y * T(0.5) + T(0.5), z * T(0.5) + T(0.5));
}
// floatpacking.glsl:46-66
template <class T> V3<T> unpack3sn(uint32_t u) {
  T a = T(u & 0x3FFu) / T(1023.0);
  T b = T((u >> 10) & 0x3FFu) / T(1023.0);
  T c = (T((u >> 20) & 0x3FFu) - T(1.5)) / T(1021.0);
  return v3<T>(a * T(2.0) - T(1.0), b * T(2.0) - T(1.0), c * T(2.0) - T(1.0));
}

uint32_t pack_unorm8(float c) {
  float v = std::min(std::max(c, 0.0f), 1.0f) * 255.0f;
  return (uint32_t)std::nearbyint(v);
}

// IEEE binary32 -> binary16, round to nearest even (textbook bit method).
uint16_t f32_to_f16(float f) {
  uint32_t x = f2u(f);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t absx = x & 0x7FFFFFFFu;
  if (absx >= 0x7F800000u) {  // inf / nan
    return (uint16_t)(sign | 0x7C00u | (absx > 0x7F800000u ? 0x200u : 0u));
  }
  if (absx >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);  // rounds to inf
  if (absx < 0x38800000u) {                                     // subnormal half or zero
    if (absx < 0x33000000u) return (uint16_t)sign;              // < 2^-25 -> 0
    uint32_t e = absx >> 23;
    uint32_t m = (absx & 0x7FFFFFu) | 0x800000u;
    uint32_t shift = 126u - e;  // value = m * 2^(e-150); half subnormal unit 2^-24
    // result = m >> (shift - 0) relative to 2^-24 unit: m*2^(e-150) / 2^-24 = m >> (126 - e)
    uint32_t q = m >> shift;
    uint32_t rem = m & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t e = (absx >> 23) - 112u;  // rebias 127 -> 15
  uint32_t m = absx & 0x7FFFFFu;
  uint32_t q = (e << 10) | (m >> 13);
  uint32_t rem = m & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
  return (uint16_t)(sign | q);
}

float f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1Fu;
  uint32_t m = h & 0x3FFu;
  if (e == 0) {
    float v = std::ldexp((float)m, -24);
    return sign ? -v : v;
  }
  if (e == 31) return u2f(sign | 0x7F800000u | (m << 13));
  return u2f(sign | ((e + 112u) << 23) | (m << 13));
}

// floatpacking.glsl:68-76 packHalf2x16 / unpackHalf2x16
template <class T> uint32_t pack_half2(T x, T y) {
  return (uint32_t)f32_to_f16((float)x) | ((uint32_t)f32_to_f16((float)y) << 16);
}
template <class T> void unpack_half2(uint32_t u, T& x, T& y) {
  x = (T)f16_to_f32((uint16_t)(u & 0xFFFFu));
  y = (T)f16_to_f32((uint16_t)(u >> 16));
}

// ----------------------------------------------------------------- random
// random.glsl:34-41
uint32_t hash1(uint32_t x) {
  x += (x << 10u);
  x ^= (x >> 6u);
  x += (x << 3u);
  x ^= (x >> 11u);
  x += (x << 15u);
  return x;
}
// random.glsl:45
uint32_t hash3(uint32_t x, uint32_t y, uint32_t z) { return hash1(x ^ hash1(y) ^ hash1(z)); }
// random.glsl:50-59
float float_construct(uint32_t m) {
  m &= 0x007FFFFFu;
  m |= 0x3F800000u;
  return u2f(m) - 1.0f;
}

// ------------------------------------------------------------- ism_utils
int ism_indices_1d(int ism_count) {
  // ism.comp:39 int(pow(2, ceil(log2(ismCount) / 2))): smallest 2^k with 4^k >= count
  int k = 0;
  while ((1LL << (2 * k)) < (long long)ism_count) ++k;
  return 1 << k;
}

// ism_utils.glsl:5-13 lookAtRH; rows of the returned matrix are s, u, -f.
template <class T> void look_at_rh(V3<T> f, V3<T>& s, V3<T>& u) {
  V3<T> up = v3<T>(T(0.0), T(1.0), T(0.01f));
  s = normalize(cross(f, up));
  u = cross(s, f);
}

// ism_utils.glsl:15-40 paraboloid_project
template <class T>
V3<T> paraboloid_project(V3<T> rel, T dist, V3<T> n, T z_far, int ism_index, int n1d,
                         bool preserve_sign) {
  V3<T> s, u;
  look_at_rh(n, s, u);
  V3<T> mf = v3<T>(-n.x, -n.y, -n.z);
  V3<T> v = v3<T>(dot(s, rel), dot(u, rel), dot(mf, rel));
  T sign_v = v.z > T(0) ? T(1) : (v.z < T(0) ? T(-1) : T(0));
  v.x = v.x / dist;
  v.y = v.y / dist;
  v.z = v.z / dist;
  v.z = T(1.0) - v.z;
  v.x = v.x / v.z;
  v.y = v.y / v.z;
  v.z = dist / z_far;
  if (preserve_sign) v.z = v.z * -sign_v;
  v.x = v.x + T(1.0);
  v.y = v.y + T(1.0);
  v.x = v.x / T(2.0);
  v.y = v.y / T(2.0);
  int ty = ism_index / n1d;  // int(ismIndex / ismIndices1d), exact for index < 2^24
  int tx = ism_index - ty * n1d;
  v.x = v.x + T(tx);
  v.y = v.y + T(ty);
  v.x = v.x / T(n1d);
  v.y = v.y / T(n1d);
  return v;
}

// ---------------------------------------------------------- VPL extract
// vpl_processor.comp:41-72 with VPLProcessor.cpp:73-101 host setup.
template <class T>
void vpl_extract(const float* diffuse, const float* normal, const float* depth, int w, int h,
                 const float* m, float light_intensity, const int32_t* shuffle, int n, T* pos,
                 T* nrm, uint32_t* packed, T* color) {
  uint32_t s1d = 0;
  while ((uint64_t)s1d * s1d < (uint64_t)n) ++s1d;  // uint(ceil(sqrt(n))), n <= 2^20
  uint32_t total = (uint32_t)n;
  uint32_t axis_x = s1d, axis_y = total / s1d;
  for (uint32_t i = 0; i < (uint32_t)n; ++i) {
    uint32_t tx = (i % axis_x) * (uint32_t)w / axis_x;
    uint32_t ty = (i / axis_x) * (uint32_t)h / axis_y;
    T uvx = T((float)tx) / T((float)w);
    T uvy = T((float)ty) / T((float)h);
    T dif[3] = {T(0), T(0), T(0)}, nr[3] = {T(0), T(0), T(0)}, d = T(0);
    if (tx < (uint32_t)w && ty < (uint32_t)h) {
      size_t p = (size_t)ty * w + tx;
      for (int c = 0; c < 3; ++c) {
        dif[c] = T(diffuse[p * 3 + c]);
        nr[c] = T(normal[p * 3 + c]);
      }
      d = T(depth[p]);
    }
    T nx = nr[0] * T(2.0) - T(1.0), ny = nr[1] * T(2.0) - T(1.0), nz = nr[2] * T(2.0) - T(1.0);
    // biasedLightViewProjectionInverseMatrix * vec4(uv, depth, 1), column-major m[col*4+row]
    T r[4];
    for (int row = 0; row < 4; ++row) {
      // depth term first: the depth and constant columns cancel for a
      // perspective light, so they are fused before the x/y terms are added
      T acc = madd(T(m[2 * 4 + row]), d, T(m[3 * 4 + row]));
      acc = madd(T(m[1 * 4 + row]), uvy, acc);
      acc = madd(T(m[0 * 4 + row]), uvx, acc);
      r[row] = acc;
    }
    uint32_t dst = shuffle ? (uint32_t)shuffle[i] : i;
    pos[dst * 3 + 0] = r[0] / r[3];
    pos[dst * 3 + 1] = r[1] / r[3];
    pos[dst * 3 + 2] = r[2] / r[3];
    T li = T(light_intensity);
    color[dst * 3 + 0] = dif[0] * li;
    color[dst * 3 + 1] = dif[1] * li;
    color[dst * 3 + 2] = dif[2] * li;
    uint32_t pk = pack3sn<T>(nx, ny, nz);
    V3<T> un = unpack3sn<T>(pk);
    packed[dst] = pk;
    nrm[dst * 3 + 0] = un.x;
    nrm[dst * 3 + 1] = un.y;
    nrm[dst * 3 + 2] = un.z;
  }
}

// ------------------------------------------------------------- ISM splat
// ism.comp:54-139 (usePushPull branch). The VPL range [start, start + count)
// and the two flags follow ism.comp:31-40,56-66,124: with scale_isms the
// atlas holds `count` ISMs and VPL v splats into ISM v - start (indices
// outside the atlas's grid write nothing or, through the truncating
// ivec2(), the first row / column, as GL's invalid image atomics and int()
// do); with points_only the `count` bins take their 16 candidates from the
// range only, ((b + i) mod n_vpl) mod count + start.
// The atlas is a per-texel minimum, so the bins may run on `threads`
// threads in any order (atomic min) with the same result.
template <class T>
void ism_splat(const float* points, const uint32_t* bin_off, int n_bins, const float* vpos,
               const float* vnrm, int n_vpl, int atlas_side, float z_far, uint32_t* atlas, int threads = 1,
               int range_start = 0, int range_count = -1, bool scale_isms = false, bool points_only = false) {
  const int max_test = 16, max_collect = 4;
  if (range_count < 0) range_count = n_vpl;
  std::fill(atlas, atlas + (size_t)atlas_side * atlas_side, 0xFFFFFFFFu);
  const int n1d = ism_indices_1d(scale_isms ? range_count : n_vpl);
  const T zf = T(z_far);
  const bool atomic = threads > 1;
  parallel_rows(n_bins, threads, [&](int b) {
    int ids[max_test];
    V3<T> cp[max_test], cn[max_test];
    for (int i = 0; i < max_test; ++i) {
      int idx = (b + i) % n_vpl;  // index %= totalVplCount
      if (points_only) idx = idx % range_count + range_start;
      ids[i] = scale_isms ? idx - range_start : idx;  // the ISM index
      cp[i] = v3<T>(T(vpos[idx * 3]), T(vpos[idx * 3 + 1]), T(vpos[idx * 3 + 2]));
      cn[i] = v3<T>(T(vnrm[idx * 3]), T(vnrm[idx * 3 + 1]), T(vnrm[idx * 3 + 2]));
    }
    for (uint32_t p = bin_off[b]; p < bin_off[b + 1]; ++p) {
      const float* rec = points + (size_t)p * 4;
      V3<T> pos = v3<T>(T(rec[0]), T(rec[1]), T(rec[2]));
      uint32_t nr = f2u(rec[3]);
      T ux = T((float)(nr & 0xFFu)) / T(255.0);
      T uy = T((float)((nr >> 8) & 0xFFu)) / T(255.0);
      T uz = T((float)((nr >> 16) & 0xFFu)) / T(255.0);
      T uw = T((float)(nr >> 24)) / T(255.0);
      V3<T> pn = v3<T>(ux * T(2.0) - T(1.0), uy * T(2.0) - T(1.0), uz * T(2.0) - T(1.0));
      T radius = uw * T(25.0);
      int used[max_collect];
      int found = 0;
      for (int i = 0; i < max_test; ++i) {
        V3<T> rel = sub(pos, cp[i]);
        V3<T> nrel = v3<T>(-rel.x, -rel.y, -rel.z);
        bool cull = dot(cn[i], rel) < T(0) || dot(pn, nrel) < T(0);
        if (!cull && found < max_collect) used[found++] = i;
      }
      for (int k = 0; k < found; ++k) {
        int li = used[k];
        V3<T> rel = sub(pos, cp[li]);
        T dist = length(rel);
        V3<T> v = paraboloid_project<T>(rel, dist, cn[li], zf, ids[li], n1d, true);
        T vx = v.x * T(atlas_side);
        T vy = v.y * T(atlas_side);
        T vz = v.z * T(16777216.0);
        uint32_t val = to_uint(vz) << 8;
        val |= to_uint(radius * T(10.0) / T(2.0));  // sqrt(float(maxVplCollectCount)) == 2
        // ivec2(v.xy): truncation toward zero
        long long tx = (long long)std::trunc(vx), ty = (long long)std::trunc(vy);
        if (tx < 0 || ty < 0 || tx >= atlas_side || ty >= atlas_side) continue;
        store_min(&atlas[(size_t)ty * atlas_side + tx], val, atomic);
      }
    }
  });
}

// ------------------------------------------------------------- pull-push
template <class T> struct Texel {
  T depth, max_depth, radius;
  uint32_t disp;  // packHalf2x16
};

// pull.comp:20-126, one level. lvl = level written (1..L).
template <class T>
void pull_level(const uint32_t* atlas, const std::vector<Texel<T>>* prev, std::vector<Texel<T>>& out,
                int out_side, int lvl, int ism_side, T z_far, int threads) {
  static const int offs[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
  const int in_side = out_side * 2;
  const T scale = std::ldexp(T(1.0), -lvl);
  out.assign((size_t)out_side * out_side, Texel<T>{T(1), T(0), T(0), 0u});
  parallel_rows(out_side, threads, [&](int oy) {
    for (int ox = 0; ox < out_side; ++ox) {
      T ds[4], mds[4], rs[4], dx[4], dy[4];
      bool valid[4];
      for (int i = 0; i < 4; ++i) {
        int ix = ox * 2 + offs[i][0], iy = oy * 2 + offs[i][1];
        T depth, max_depth, radius, disp_x = T(0), disp_y = T(0);
        if (lvl == 1) {
          uint32_t s = atlas[(size_t)iy * in_side + ix];
          depth = T((float)(s >> 8)) / T(16777216.0);
          radius = T((float)(s & 0xFFu)) / T(10.0);
          max_depth = depth + radius * T(2.0) / z_far * T(0.6f);
          T dist = depth * z_far;
          radius = radius / dist / T(3.14f) * T(ism_side);
          radius = radius * T(1.3f);
          radius = std::min(radius, T(15.0));
        } else {
          const Texel<T>& t = (*prev)[(size_t)iy * in_side + ix];
          depth = t.depth;
          max_depth = t.max_depth;
          unpack_half2<T>(t.disp, disp_x, disp_y);
          radius = t.radius;
        }
        T ndx = (T(ix) + T(0.5) + disp_x) / T(2.0) - (T(ox) + T(0.5));
        T ndy = (T(iy) + T(0.5) + disp_y) / T(2.0) - (T(oy) + T(0.5));
        T dist = std::sqrt(ndx * ndx + ndy * ndy);
        T scaled = radius * scale;
        ds[i] = depth;
        mds[i] = max_depth;
        rs[i] = radius;
        dx[i] = ndx;
        dy[i] = ndy;
        valid[i] = depth != T(1.0) && dist <= scaled;
      }
      T minimum = T(INFINITY);
      T max_depth = T(0);
      bool any = false;
      for (int i = 0; i < 4; ++i) {
        if (!valid[i]) continue;
        any = true;
        minimum = std::min(ds[i], minimum);
        if (minimum == ds[i]) max_depth = mds[i];
      }
      if (!any) continue;  // result stays (1, 0, 0, 0)
      for (int i = 0; i < 4; ++i)
        if (ds[i] > max_depth) valid[i] = false;
      T dacc = T(0), racc = T(0), xacc = T(0), yacc = T(0), mvm = T(0);
      uint32_t nv = 0;
      for (int i = 0; i < 4; ++i) {
        if (!valid[i]) continue;
        dacc = dacc + ds[i];
        xacc = xacc + dx[i];
        yacc = yacc + dy[i];
        racc = racc + rs[i];
        mvm = std::max(mvm, mds[i]);
        nv++;
      }
      Texel<T>& o = out[(size_t)oy * out_side + ox];
      if (nv > 0) {
        T fn = T((float)nv);
        o.depth = dacc / fn;
        o.max_depth = mvm;
        o.radius = racc / fn;
        o.disp = pack_half2<T>(xacc / fn, yacc / fn);
      }
    }
  });
}

inline int log2i(int v) {
  int k = 0;
  while ((1 << k) < v) ++k;
  return k;
}

// push.comp:44-165, one level. lvl = level written (L-1..0).
// same: pull level lvl (lvl >= 1); atlas for lvl == 0. coarse: level lvl+1.
// Invocation row gy writes fine rows 2 gy - 1 and 2 gy only, so the rows of
// invocations run on `threads` threads.
template <class T>
void push_level(const uint32_t* atlas, const std::vector<Texel<T>>* same,
                const std::vector<Texel<T>>& coarse, std::vector<Texel<T>>* out,
                uint16_t* out_r16, T* out_depth, int side, int lvl, int ism_side, int threads) {
  static const int offs[4][2] = {{0, 0}, {0, 1}, {1, 1}, {1, 0}};
  static const int wx[4] = {9, 3, 1, 3};
  const int cside = side / 2;
  const int shift = log2i(ism_side) - (lvl + 1);
  const T scale = std::ldexp(T(1.0), -lvl);
  if (out) out->assign((size_t)side * side, Texel<T>{T(1), T(0), T(0), 0u});
  const int groups = side / 2 + 1;  // invocations per axis that touch [0, side)
  parallel_rows(groups, threads, [&](int gy) {
    for (int gx = 0; gx < groups; ++gx) {
      int clx = gx - 1, cly = gy - 1;
      T cd[4], cmd[4], cr[4], cdx[4], cdy[4];
      for (int i = 0; i < 4; ++i) {
        int x = clx + offs[i][0], y = cly + offs[i][1];
        if (x < 0 || y < 0 || x >= cside || y >= cside) {
          cd[i] = cmd[i] = cr[i] = cdx[i] = cdy[i] = T(0);  // imageLoad out of range = 0
          continue;
        }
        const Texel<T>& t = coarse[(size_t)y * cside + x];
        cd[i] = t.depth;
        cmd[i] = t.max_depth;
        cr[i] = t.radius;
        unpack_half2<T>(t.disp, cdx[i], cdy[i]);
      }
      for (int op = 0; op < 4; ++op) {
        int px = gx * 2 - 1 + offs[op][0], py = gy * 2 - 1 + offs[op][1];
        if (px < 0 || py < 0 || px >= side || py >= side) continue;  // imageStore no-op
        int w[4];
        for (int i = 0; i < 4; ++i) w[i] = wx[(i - op + 4) % 4];
        int otx = px / 2, oty = py / 2;
        for (int i = 0; i < 4; ++i) {
          int ix = clx + offs[i][0], iy = cly + offs[i][1];
          if ((otx >> shift) != (ix >> shift) || (oty >> shift) != (iy >> shift)) w[i] = 0;
        }
        for (int i = 0; i < 4; ++i)
          if (cd[i] == T(1.0)) w[i] = 0;
        T ddx[4], ddy[4];
        for (int i = 0; i < 4; ++i) {
          T ctx = (T(clx + offs[i][0]) + T(0.5)) * T(2.0);
          T cty = (T(cly + offs[i][1]) + T(0.5)) * T(2.0);
          T ttx = T(px) + T(0.5), tty = T(py) + T(0.5);
          ddx[i] = ctx - ttx + cdx[i] * T(2.0);
          ddy[i] = cty - tty + cdy[i] * T(2.0);
          T dist = std::sqrt(ddx[i] * ddx[i] + ddy[i] * ddy[i]);
          T radius = cr[i] * scale;
          if (dist > radius) w[i] = 0;
        }
        T minimum = T(9001.0);
        T max_depth = T(0);
        for (int i = 0; i < 4; ++i) {
          if (w[i] == 0) continue;
          minimum = std::min(cd[i], minimum);
          if (minimum == cd[i]) max_depth = cmd[i];
        }
        int wacc = 0;
        for (int i = 0; i < 4; ++i) wacc += w[i];
        if (wacc > 0) {
          for (int i = 0; i < 4; ++i)
            if (cd[i] > max_depth) w[i] = 0;
        }
        T dacc = T(0), macc = T(0), racc = T(0), xacc = T(0), yacc = T(0);
        wacc = 0;
        for (int i = 0; i < 4; ++i) {
          if (w[i] == 0) continue;  // zero-weight terms add +-0: skipping is exact
          T fw = T((float)w[i]);
          dacc = dacc + cd[i] * fw;
          macc = macc + cmd[i] * fw;
          racc = racc + cr[i] * fw;
          xacc = xacc + ddx[i] * fw;
          yacc = yacc + ddy[i] * fw;
          wacc += w[i];
        }
        Texel<T> res{T(0), T(0), T(0), 0u};
        if (wacc > 0) {
          T fw = T((float)wacc);
          res.depth = dacc / fw;
          res.max_depth = macc / fw;
          res.radius = racc / fw;
          res.disp = pack_half2<T>(xacc / fw, yacc / fw);
        }
        Texel<T> orig;
        if (lvl == 0) {
          uint32_t s = atlas[(size_t)py * side + px];
          orig = Texel<T>{T((float)(s >> 8)) / T(16777216.0), T(0), T(0), 0u};
        } else {
          orig = (*same)[(size_t)py * side + px];
        }
        bool invalid = orig.depth == T(1.0);
        bool occluded = false;
        for (int i = 0; i < 4; ++i) occluded = occluded || (w[i] > 0 && orig.depth > cmd[i]);
        if (wacc <= 0 || (!invalid && !occluded)) res = orig;
        if (lvl == 0) {
          T r = std::min(std::max(res.depth, T(0)), T(1));
          out_r16[(size_t)py * side + px] = (uint16_t)std::nearbyint(r * T(65535.0));
          if (out_depth) out_depth[(size_t)py * side + px] = res.depth;
        } else {
          (*out)[(size_t)py * side + px] = res;
        }
      }
    }
  });
}

template <class T>
void pullpush(const uint32_t* atlas, int atlas_side, int ism_side, int levels, float z_far,
              uint16_t* out_r16, T* out_depth, int threads = 1) {
  std::vector<std::vector<Texel<T>>> pull(levels + 1);
  for (int l = 1; l <= levels; ++l)
    pull_level<T>(atlas, l > 1 ? &pull[l - 1] : nullptr, pull[l], atlas_side >> l, l, ism_side,
                  T(z_far), threads);
  std::vector<Texel<T>> coarse = pull[levels], cur;
  for (int l = levels - 1; l >= 0; --l) {
    if (l > 0) {
      push_level<T>(atlas, &pull[l], coarse, &cur, nullptr, nullptr, atlas_side >> l, l, ism_side, threads);
      coarse.swap(cur);
    } else {
      push_level<T>(atlas, nullptr, coarse, nullptr, out_r16, out_depth, atlas_side, 0, ism_side, threads);
    }
  }
}

// ---------------------------------------------------------- final gather
template <class T> T smoothstep(T e0, T e1, T x) {
  T t = std::min(std::max((x - e0) / (e1 - e0), T(0)), T(1));
  return t * t * (T(3) - T(2) * t);
}

// ism_offset / ism_count: SCALE_ISMS's ismIndexOffset and ismCount
// (final_gathering.comp:43-45): VPL j samples ISM j - ism_offset of a grid
// laid out for ism_count ISMs (ism_count < 0: n_vpl).
template <class T>
void final_gather_rows(const float* depth, const float* normal, int w, int h, int y0, int y1,
                       const float* m, const float* vpos, const float* vnrm, const float* vcol,
                       int n_vpl, int vpl_start, int vpl_end, const uint16_t* ism, int atlas_side,
                       float z_far, float gi_intensity, float clamp_value, int interleave,
                       int shadowing, int threads, T* out, uint64_t* lit_pairs, int ism_offset = 0,
                       int ism_count = -1) {
  const int classes = interleave * interleave;
  const int per_class = n_vpl / classes;
  const int n1d = ism_indices_1d(ism_count < 0 ? n_vpl : ism_count);
  const T zf = T(z_far);
  const T bias = T(0.02f);
  const T vpl_count = T((float)(vpl_end - vpl_start));
  std::atomic<uint64_t> lit{0};
  std::atomic<int> next_row{y0};
  auto worker = [&]() {
    uint64_t my_lit = 0;
    for (;;) {
      int y = next_row.fetch_add(1);
      if (y >= y1) break;
      for (int x = 0; x < w; ++x) {
        size_t p = (size_t)y * w + x;
        T uvx = T((float)x) / T((float)w), uvy = T((float)y) / T((float)h);
        T ndc[4] = {uvx * T(2.0) - T(1.0), uvy * T(2.0) - T(1.0), T(depth[p]) * T(2.0) - T(1.0),
                    T(1.0) * T(2.0) - T(1.0)};
        // vec4(v_uv, depthSample, 1.0) * 2.0 - 1.0: the w component becomes 1.0
        T r[4];
        for (int row = 0; row < 4; ++row) {
          T acc = madd(T(m[2 * 4 + row]), ndc[2], T(m[3 * 4 + row]) * ndc[3]);
          acc = madd(T(m[1 * 4 + row]), ndc[1], acc);
          r[row] = madd(T(m[0 * 4 + row]), ndc[0], acc);
        }
        V3<T> frag = v3<T>(r[0] / r[3], r[1] / r[3], r[2] / r[3]);
        V3<T> fn = v3<T>(T(normal[p * 3]) * T(2.0) - T(1.0), T(normal[p * 3 + 1]) * T(2.0) - T(1.0),
                         T(normal[p * 3 + 2]) * T(2.0) - T(1.0));
        int cls = (x % interleave) + (y % interleave) * interleave;
        int start = cls * per_class;
        T ar = T(0), ag = T(0), ab = T(0);
        for (int j = start; j < start + per_class; ++j) {
          V3<T> vp = v3<T>(T(vpos[j * 3]), T(vpos[j * 3 + 1]), T(vpos[j * 3 + 2]));
          V3<T> vn = v3<T>(T(vnrm[j * 3]), T(vnrm[j * 3 + 1]), T(vnrm[j * 3 + 2]));
          V3<T> diff = sub(frag, vp);
          T dist = length(diff);
          V3<T> nd = v3<T>(diff.x / dist, diff.y / dist, diff.z / dist);
          V3<T> mnd = v3<T>(-nd.x, -nd.y, -nd.z);
          T angle = std::max(T(0), dot(vn, nd)) * std::max(T(0), dot(fn, mnd));
          if (!(angle > T(0))) continue;
          ++my_lit;
          T d2 = dist * dist;
          T att = T(1) / (d2 * d2);
          T g = std::min(angle * att, T(clamp_value));
          if (shadowing) {
            V3<T> v = paraboloid_project<T>(diff, dist, vn, zf, j - ism_offset, n1d, false);
            T fx = std::floor(v.x * T(atlas_side)), fy = std::floor(v.y * T(atlas_side));
            long long tx = ((long long)fx % atlas_side + atlas_side) % atlas_side;
            long long ty = ((long long)fy % atlas_side + atlas_side) % atlas_side;
            T occ = T((float)ism[(size_t)ty * atlas_side + tx]) / T(65535.0);
            T sv = v.z - occ;
            sv = smoothstep(T(1.0) - bias, T(1.0), T(1) - sv);
            g = g * sv;
          }
          ar = ar + T(vcol[j * 3]) * g;
          ag = ag + T(vcol[j * 3 + 1]) * g;
          ab = ab + T(vcol[j * 3 + 2]) * g;
        }
        T ip = T((float)classes);
        size_t o = ((size_t)(y - y0) * w + x) * 3;
        out[o] = ar * T(gi_intensity) / vpl_count * ip;
        out[o + 1] = ag * T(gi_intensity) / vpl_count * ip;
        out[o + 2] = ab * T(gi_intensity) / vpl_count * ip;
      }
    }
    lit.fetch_add(my_lit);
  };
  int nt = std::max(1, threads);
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
  if (lit_pairs) *lit_pairs = lit.load();
}

}  // namespace

// ------------------------------------------------------ geometry-aware blur
// gi/gi_blur.frag:24-72 run as GIStage::blur does (GIStage.cpp:308-344):
// pass x (DIRECTION (1,0)) into a temporary, then pass y (DIRECTION (0,1))
// of the temporary. Both passes read the same face normals and depth.
// linearDepth (reprojection.glsl:6-13) uses P[2][2] and P[3][2]
// (column-major m[10], m[14]). Out-of-range texelFetch reads 0.
template <class T> T blur_linear_depth(T d, T A, T B) {
  T zn = T(2.0) * d - T(1.0);
  return -B / (A + zn);
}

template <class T>
void gi_blur_pass(const T* gi, const float* normal, const float* depth, int w, int h, const float* proj,
                  int dx, int dy, T* out) {
  const T A = T(proj[10]), B = T(proj[14]);
  const int radius = 3;
  for (int y = 0; y < h; ++y) {
    for (int x = 0; x < w; ++x) {
      const size_t c = (size_t)y * w + x;
      const T d = blur_linear_depth<T>(T(depth[c]), A, B);
      const T nx = T(normal[c * 3]) * T(2.0) - T(1.0), ny = T(normal[c * 3 + 1]) * T(2.0) - T(1.0),
              nz = T(normal[c * 3 + 2]) * T(2.0) - T(1.0);
      T acc[3] = {gi[c * 3], gi[c * 3 + 1], gi[c * 3 + 2]};
      T fsum = T(1.0);
      for (int side = -1; side <= 1; side += 2) {  // the -DIRECTION loop, then +DIRECTION
        for (int i = 1; i <= radius; ++i) {
          const int sx = x + side * dx * i, sy = y + side * dy * i;
          const bool in = sx >= 0 && sy >= 0 && sx < w && sy < h;
          const size_t q = (size_t)sy * w + sx;
          const T g0 = in ? gi[q * 3] : T(0), g1 = in ? gi[q * 3 + 1] : T(0), g2 = in ? gi[q * 3 + 2] : T(0);
          const T mx = (in ? T(normal[q * 3]) : T(0)) * T(2.0) - T(1.0);
          const T my = (in ? T(normal[q * 3 + 1]) : T(0)) * T(2.0) - T(1.0);
          const T mz = (in ? T(normal[q * 3 + 2]) : T(0)) * T(2.0) - T(1.0);
          const T ds = blur_linear_depth<T>(in ? T(depth[q]) : T(0), A, B);
          const T nf = T(1.0) - std::max(T(0), nx * mx + ny * my + nz * mz);
          const T dd = (ds - d) / T(1.0);
          const T nd = nf * T(5.0);
          // weight(r, depthDiff, normalDiff): max(0, (radius - r + 1) / 4 - dd^2 - nd^2)
          const T factor = std::max(T(0), (T(radius) - T(i) + T(1.0)) / T(4.0) - dd * dd - nd * nd);
          fsum = fsum + factor;
          acc[0] = acc[0] + g0 * factor;
          acc[1] = acc[1] + g1 * factor;
          acc[2] = acc[2] + g2 * factor;
        }
      }
      for (int k = 0; k < 3; ++k) out[c * 3 + k] = acc[k] / fsum;
    }
  }
}

template <class T>
void gi_blur(const float* gi, const float* normal, const float* depth, int w, int h, const float* proj, T* out) {
  std::vector<T> in((size_t)w * h * 3), tmp((size_t)w * h * 3);
  for (size_t i = 0; i < in.size(); ++i) in[i] = T(gi[i]);
  gi_blur_pass<T>(in.data(), normal, depth, w, h, proj, 1, 0, tmp.data());
  gi_blur_pass<T>(tmp.data(), normal, depth, w, h, proj, 0, 1, out);
}

// ------------------------------------------------------ deferred composite
// deferredshading.frag:47-81 with shadowmapping.glsl:6-34 (VSM) and
// srgb_utils.glsl (gamma 2.2), as DeferredShadingStage::process binds it
// (DeferredShadingStage.cpp:76-131). Choices:
//   * the texel (x, y) is sampled at its centre, uv = ((x+0.5)/w, (y+0.5)/h);
//     the G-buffer, GI and occlusion textures are NEAREST, so they read
//     texel (x, y);
//   * the world position is inverse(P*V) applied to the pixel-centre NDC
//     (2uv-1, 2 depth-1, 1), divided by w: the same point as
//     viewInverted * (linearDepth * viewRay) for a perspective camera;
//   * the VSM moments are sampled bilinearly (GL_LINEAR, texel centres at
//     (i+0.5)/W) with CLAMP_TO_BORDER, border (0, 0) (RasterizationStage.cpp:97-99);
//   * pow(x, y) = exp2(y log2 x) with pow(0, y) = 0 for y > 0;
//   * a null specular buffer reads 0, a null occlusion buffer reads 1.
template <class T> T c_clamp01(T x) { return std::min(std::max(x, T(0)), T(1)); }
template <class T> T c_smoothstep(T e0, T e1, T x) {
  const T t = c_clamp01((x - e0) / (e1 - e0));
  return t * t * (T(3) - T(2) * t);
}
template <class T> T c_pow(T x, T y) { return x > T(0) ? std::pow(x, y) : T(0); }

template <class T>
T c_vsm(T mx, T my, T compare) {
  const T p = c_smoothstep(compare - T(0.001f), compare, mx);
  const T variance = std::max(my - mx * mx, T(0.001f));
  const T d = compare - mx;
  const T p_max = c_clamp01((variance / (variance + d * d) - T(0.1f)) / (T(1.0) - T(0.1f)));  // linstep(0.1, 1, .)
  return c_clamp01(std::max(p, p_max));
}

template <class T>
void c_moments(const float* vsm, int W, int H, T u, T v, T* mx, T* my) {
  const T s = u * T(W) - T(0.5), t = v * T(H) - T(0.5);
  const T fs = std::floor(s), ft = std::floor(t);
  T m[2] = {T(0), T(0)};
  *mx = *my = T(0);
  if (!(std::fabs(fs) < T(1e9) && std::fabs(ft) < T(1e9))) return;  // far outside (or w = 0): border
  const int i0 = (int)fs, j0 = (int)ft;
  const T a = s - fs, b = t - ft;
  for (int k = 0; k < 4; ++k) {
    const int i = i0 + (k & 1), j = j0 + (k >> 1);
    const T w = ((k & 1) ? a : T(1) - a) * ((k >> 1) ? b : T(1) - b);
    if (i < 0 || j < 0 || i >= W || j >= H) continue;  // border (0, 0)
    const size_t q = ((size_t)j * W + i) * 2;
    m[0] += w * T(vsm[q]);
    m[1] += w * T(vsm[q + 1]);
  }
  *mx = m[0];
  *my = m[1];
}

template <class T>
void deferred_composite(const float* depth, const float* normal, const float* diffuse, const float* specular,
                        const float* gi, const float* occlusion, int w, int h, const float* vsm, int vw, int vh,
                        const float* vp_inv, const float* blvp, const float* eye, const float* light_pos,
                        const float* inv_light_dir, float light_intensity, float exposure, T* out) {
  const T gamma = T(2.2f);
  for (int y = 0; y < h; ++y) {
    for (int x = 0; x < w; ++x) {
      const size_t c = (size_t)y * w + x;
      const T u = (T(x) + T(0.5)) / T(w), v = (T(y) + T(0.5)) / T(h);
      const T ndc[4] = {u * T(2) - T(1), v * T(2) - T(1), T(depth[c]) * T(2) - T(1), T(1)};
      T r[4];
      for (int row = 0; row < 4; ++row) {
        r[row] = T(0);
        for (int col = 0; col < 4; ++col) r[row] += T(vp_inv[col * 4 + row]) * ndc[col];
      }
      const T wx = r[0] / r[3], wy = r[1] / r[3], wz = r[2] / r[3];
      const T nx = T(normal[c * 3]) * T(2) - T(1), ny = T(normal[c * 3 + 1]) * T(2) - T(1),
              nz = T(normal[c * 3 + 2]) * T(2) - T(1);
      const T lx = T(inv_light_dir[0]), ly = T(inv_light_dir[1]), lz = T(inv_light_dir[2]);
      T vx = T(eye[0]) - wx, vy = T(eye[1]) - wy, vz = T(eye[2]) - wz;
      const T vl = std::sqrt(vx * vx + vy * vy + vz * vz);
      vx /= vl, vy /= vl, vz /= vl;
      T hx = lx + vx, hy = ly + vy, hz = lz + vz;
      const T hl = std::sqrt(hx * hx + hy * hy + hz * hz);
      hx /= hl, hy /= hl, hz /= hl;
      const T ndotl = nx * lx + ny * ly + nz * lz;
      const T ndoth = nx * hx + ny * hy + nz * hz;
      T sc[4];
      const T wp[4] = {wx, wy, wz, T(1)};
      for (int row = 0; row < 4; ++row) {
        sc[row] = T(0);
        for (int col = 0; col < 4; ++col) sc[row] += T(blvp[col * 4 + row]) * wp[col];
      }
      T mx, my;
      c_moments<T>(vsm, vw, vh, sc[0] / sc[3], sc[1] / sc[3], &mx, &my);
      const T dlx = wx - T(light_pos[0]), dly = wy - T(light_pos[1]), dlz = wz - T(light_pos[2]);
      const T dist = std::sqrt(dlx * dlx + dly * dly + dlz * dlz);
      T shadow = c_vsm(mx, my, dist);
      shadow *= sc[3] > T(0) ? T(1) : T(0);  // step(0, sign(w))
      const T occ = occlusion ? T(occlusion[c]) : T(1);
      const T spec_f = c_pow(std::max(T(0), ndoth), T(20));
      for (int k = 0; k < 3; ++k) {
        const T dif = c_pow(T(diffuse[c * 3 + k]), gamma);
        const T spc = specular ? c_pow(T(specular[c * 3 + k]), gamma) : T(0);
        const T ambient = T(gi[c * 3 + k]) * dif * occ;
        const T dterm = dif * (std::max(T(0), ndotl) * shadow) * T(light_intensity);
        const T sterm = T(0.75f) * spc * spec_f * shadow;
        T col = (ambient + dterm + sterm) * T(exposure);
        col = col / (T(1) + col);
        out[c * 3 + k] = c_pow(col, T(1) / gamma);
      }
    }
  }
}

extern "C" {

void oracle_deferred_composite_f64(const float* depth, const float* normal, const float* diffuse,
                                   const float* specular, const float* gi, const float* occlusion, int w, int h,
                                   const float* vsm, int vw, int vh, const float* vp_inv, const float* blvp,
                                   const float* eye, const float* light_pos, const float* inv_light_dir,
                                   float light_intensity, float exposure, double* out) {
  deferred_composite<double>(depth, normal, diffuse, specular, gi, occlusion, w, h, vsm, vw, vh, vp_inv, blvp, eye,
                             light_pos, inv_light_dir, light_intensity, exposure, out);
}

uint32_t oracle_pack3sn_f32(float x, float y, float z) { return pack3sn<float>(x, y, z); }
uint32_t oracle_pack3sn_f64(double x, double y, double z) { return pack3sn<double>(x, y, z); }
void oracle_unpack3sn_f32(uint32_t bits, float* o) {
  V3<float> v = unpack3sn<float>(bits);
  o[0] = v.x; o[1] = v.y; o[2] = v.z;
}
void oracle_unpack3sn_f64(uint32_t bits, double* o) {
  V3<double> v = unpack3sn<double>(bits);
  o[0] = v.x; o[1] = v.y; o[2] = v.z;
}
uint32_t oracle_pack_unorm4x8(float x, float y, float z, float w) {
  return pack_unorm8(x) | (pack_unorm8(y) << 8) | (pack_unorm8(z) << 16) | (pack_unorm8(w) << 24);
}
void oracle_unpack_unorm4x8(uint32_t b, float* o) {
  for (int i = 0; i < 4; ++i) o[i] = (float)((b >> (8 * i)) & 0xFFu) / 255.0f;
}
uint16_t oracle_f32_to_f16(float x) { return f32_to_f16(x); }
float oracle_f16_to_f32(uint16_t h) { return f16_to_f32(h); }

uint32_t oracle_hash1(uint32_t x) { return hash1(x); }
uint32_t oracle_hash3(uint32_t x, uint32_t y, uint32_t z) { return hash3(x, y, z); }
float oracle_random3(float x, float y, float z) { return float_construct(hash3(f2u(x), f2u(y), f2u(z))); }

void oracle_look_at_rh_f64(const double* n, double* s3, double* u3) {
  V3<double> s, u;
  look_at_rh<double>(v3(n[0], n[1], n[2]), s, u);
  s3[0] = s.x; s3[1] = s.y; s3[2] = s.z;
  u3[0] = u.x; u3[1] = u.y; u3[2] = u.z;
}
void oracle_look_at_rh_f32(const float* n, float* s3, float* u3) {
  V3<float> s, u;
  look_at_rh<float>(v3(n[0], n[1], n[2]), s, u);
  s3[0] = s.x; s3[1] = s.y; s3[2] = s.z;
  u3[0] = u.x; u3[1] = u.y; u3[2] = u.z;
}
void oracle_paraboloid_project_f64(const double* r, double dist, const double* n, double z_far,
                                   int idx, int n1d, int ps, double* o) {
  V3<double> v = paraboloid_project<double>(v3(r[0], r[1], r[2]), dist, v3(n[0], n[1], n[2]), z_far,
                                            idx, n1d, ps != 0);
  o[0] = v.x; o[1] = v.y; o[2] = v.z;
}
void oracle_paraboloid_project_f32(const float* r, float dist, const float* n, float z_far, int idx,
                                   int n1d, int ps, float* o) {
  V3<float> v = paraboloid_project<float>(v3(r[0], r[1], r[2]), dist, v3(n[0], n[1], n[2]), z_far,
                                          idx, n1d, ps != 0);
  o[0] = v.x; o[1] = v.y; o[2] = v.z;
}
int oracle_ism_indices_1d(int c) { return ism_indices_1d(c); }

void oracle_vpl_extract_f32(const float* d, const float* nm, const float* dp, int w, int h,
                            const float* m, float li, const int32_t* sh, int n, float* pos,
                            float* nrm, uint32_t* pk, float* col) {
  vpl_extract<float>(d, nm, dp, w, h, m, li, sh, n, pos, nrm, pk, col);
}
void oracle_vpl_extract_f64(const float* d, const float* nm, const float* dp, int w, int h,
                            const float* m, float li, const int32_t* sh, int n, double* pos,
                            double* nrm, uint32_t* pk, double* col) {
  vpl_extract<double>(d, nm, dp, w, h, m, li, sh, n, pos, nrm, pk, col);
}

void oracle_ism_splat_f32(const float* p, const uint32_t* bo, int nb, const float* vp,
                          const float* vn, int nv, int side, float zf, uint32_t* atlas) {
  ism_splat<float>(p, bo, nb, vp, vn, nv, side, zf, atlas);
}
void oracle_ism_splat_f64(const float* p, const uint32_t* bo, int nb, const float* vp,
                          const float* vn, int nv, int side, float zf, uint32_t* atlas) {
  ism_splat<double>(p, bo, nb, vp, vn, nv, side, zf, atlas);
}

void oracle_pullpush_f32(const uint32_t* a, int side, int s, int l, float zf, uint16_t* o) {
  pullpush<float>(a, side, s, l, zf, o, nullptr);
}
void oracle_ism_splat_mt_f32(const float* p, const uint32_t* bo, int nb, const float* vp, const float* vn, int nv,
                             int side, float zf, uint32_t* atlas, int threads) {
  ism_splat<float>(p, bo, nb, vp, vn, nv, side, zf, atlas, threads);
}
void oracle_pullpush_mt_f32(const uint32_t* a, int side, int s, int l, float zf, uint16_t* o, int threads) {
  pullpush<float>(a, side, s, l, zf, o, nullptr, threads);
}
void oracle_pullpush_f64(const uint32_t* a, int side, int s, int l, float zf, uint16_t* o,
                         double* od) {
  pullpush<double>(a, side, s, l, zf, o, od);
}

void oracle_final_gather_f64(const float* depth, const float* normal, int w, int h,
                             const float* vp_inv, const float* vpos, const float* vnrm,
                             const float* vcol, int n_vpl, int vs, int ve, const uint16_t* ism,
                             int side, float zf, float gi, float cl, int il, int sh, int th,
                             double* out, uint64_t* lit) {
  final_gather_rows<double>(depth, normal, w, h, 0, h, vp_inv, vpos, vnrm, vcol, n_vpl, vs, ve, ism,
                            side, zf, gi, cl, il, sh, th, out, lit);
}
void oracle_final_gather_f32(const float* depth, const float* normal, int w, int h,
                             const float* vp_inv, const float* vpos, const float* vnrm,
                             const float* vcol, int n_vpl, int vs, int ve, const uint16_t* ism,
                             int side, float zf, float gi, float cl, int il, int sh, int th,
                             float* out, uint64_t* lit) {
  final_gather_rows<float>(depth, normal, w, h, 0, h, vp_inv, vpos, vnrm, vcol, n_vpl, vs, ve, ism,
                           side, zf, gi, cl, il, sh, th, out, lit);
}
// VPL-range forms (scaleISMs / pointsOnlyIntoScaledISMs).
void oracle_ism_splat_range_f32(const float* p, const uint32_t* bo, int nb, const float* vp, const float* vn,
                                int nv, int side, float zf, uint32_t* atlas, int start, int count, int scale,
                                int points_only) {
  ism_splat<float>(p, bo, nb, vp, vn, nv, side, zf, atlas, 1, start, count, scale != 0, points_only != 0);
}
void oracle_final_gather_range_f64(const float* depth, const float* normal, int w, int h,
                                   const float* vp_inv, const float* vpos, const float* vnrm,
                                   const float* vcol, int n_vpl, int vs, int ve, const uint16_t* ism,
                                   int side, float zf, float gi, float cl, int il, int sh, int th,
                                   double* out, uint64_t* lit, int ism_offset, int ism_count) {
  final_gather_rows<double>(depth, normal, w, h, 0, h, vp_inv, vpos, vnrm, vcol, n_vpl, vs, ve, ism,
                            side, zf, gi, cl, il, sh, th, out, lit, ism_offset, ism_count);
}
void oracle_gi_blur_f32(const float* gi, const float* normal, const float* depth, int w, int h,
                        const float* proj16, float* out) {
  gi_blur<float>(gi, normal, depth, w, h, proj16, out);
}
void oracle_gi_blur_f64(const float* gi, const float* normal, const float* depth, int w, int h,
                        const float* proj16, double* out) {
  gi_blur<double>(gi, normal, depth, w, h, proj16, out);
}

void oracle_final_gather_rows_f32(const float* depth, const float* normal, int w, int h, int y0,
                                  int y1, const float* vp_inv, const float* vpos, const float* vnrm,
                                  const float* vcol, int n_vpl, int vs, int ve, const uint16_t* ism,
                                  int side, float zf, float gi, float cl, int il, int sh, int th,
                                  float* out, uint64_t* lit) {
  final_gather_rows<float>(depth, normal, w, h, y0, y1, vp_inv, vpos, vnrm, vcol, n_vpl, vs, ve,
                           ism, side, zf, gi, cl, il, sh, th, out, lit);
}

}  // extern "C"
