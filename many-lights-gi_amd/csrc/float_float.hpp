// Float-float ("double-single") arithmetic in FP32 alone, and the final
// gather's exact NEAREST texel decision built on it.
//
// A value is an unevaluated sum hi + lo of two floats with |lo| <= ulp(hi)/2,
// about 44 significant bits. The building blocks are the error-free
// transformations (Knuth's two-sum, Dekker's fast two-sum, the FMA product
// error) and Newton-corrected division and square root (the double-single
// arithmetic of Dekker 1971 and of the QD / DSFUN libraries). The operations
// are written so that no compiler contraction can change them: every product
// whose error matters goes through an explicit fmaf, and the sums are plain
// IEEE additions in a fixed order.
//
// Header-only, for the device (final gather) and the host (its CPU test).
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MLGI_FF __host__ __device__ inline
#else
#define MLGI_FF inline
#endif
// No multiply-add contraction inside these functions, whatever the including
// translation unit's setting: a product fused into a following sum would no
// longer be the rounded product the error-free transformations assume.
#if defined(__clang__)
#define MLGI_FF_EXACT _Pragma("clang fp contract(off)")
#else
#define MLGI_FF_EXACT
#endif

namespace mlgi {
namespace ff {

struct F {
  float hi, lo;
};

MLGI_FF F from(float a) { return F{a, 0.0f}; }

// a + b = s + e exactly (Knuth)
MLGI_FF F two_sum(float a, float b) {
  MLGI_FF_EXACT
  const float s = a + b;
  const float bb = s - a;
  const float e = (a - (s - bb)) + (b - bb);
  return F{s, e};
}
// a + b = s + e exactly when |a| >= |b| (Dekker)
MLGI_FF F quick_sum(float a, float b) {
  MLGI_FF_EXACT
  const float s = a + b;
  return F{s, b - (s - a)};
}
// a * b = p + e exactly (the FMA computes the product's rounding error)
MLGI_FF F two_prod(float a, float b) {
  MLGI_FF_EXACT
  const float p = a * b;
  return F{p, fmaf(a, b, -p)};
}

MLGI_FF F add(F a, F b) {
  MLGI_FF_EXACT
  F s = two_sum(a.hi, b.hi);
  const F t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = quick_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_sum(s.hi, s.lo);
}
MLGI_FF F neg(F a) { return F{-a.hi, -a.lo}; }
MLGI_FF F sub(F a, F b) { return add(a, neg(b)); }
MLGI_FF F add(F a, float b) { return add(a, from(b)); }

MLGI_FF F mul(F a, F b) {
  MLGI_FF_EXACT
  F p = two_prod(a.hi, b.hi);
  p.lo = fmaf(a.hi, b.lo, fmaf(a.lo, b.hi, p.lo));
  return quick_sum(p.hi, p.lo);
}
MLGI_FF F mul(F a, float b) {
  MLGI_FF_EXACT
  F p = two_prod(a.hi, b);
  p.lo = fmaf(a.lo, b, p.lo);
  return quick_sum(p.hi, p.lo);
}

// a / b: a first quotient, then two corrections from the exact remainders.
MLGI_FF F div(F a, F b) {
  MLGI_FF_EXACT
  const float q1 = a.hi / b.hi;
  const F r1 = sub(a, mul(b, q1));
  const float q2 = r1.hi / b.hi;
  const F r2 = sub(r1, mul(b, q2));
  const float q3 = r2.hi / b.hi;
  return add(quick_sum(q1, q2), q3);
}

// sqrt(a), a >= 0: x = sqrt(hi), then one Newton step on the exact residual.
MLGI_FF F sqrt(F a) {
  MLGI_FF_EXACT
  if (!(a.hi > 0.0f)) return F{0.0f, 0.0f};
  const float x = sqrtf(a.hi);
  const F r = sub(a, two_prod(x, x));
  return quick_sum(x, r.hi / (2.0f * x));
}

// floor of hi + lo (an integer hi is moved down when lo is negative)
MLGI_FF int floor_int(F a) {
  MLGI_FF_EXACT
  const float f = floorf(a.hi);
  return (int)f - ((f == a.hi && a.lo < 0.0f) ? 1 : 0);
}

struct F3 {
  F x, y, z;
};
MLGI_FF F dot(F3 a, F3 b) { return add(add(mul(a.x, b.x), mul(a.y, b.y)), mul(a.z, b.z)); }

// NEAREST texel (tx, ty) of a lit pair in its VPL's S x S ISM, decided with
// ~44-bit arithmetic from the FP32 inputs: the pixel (x, y) of a w x h
// G-buffer with window depth `depth`, reconstructed through the column-major
// inverse view-projection m (ndc = (x / w, y / h, depth) * 2 - 1, divided by
// w), the VPL at vp with normal vn, and the paraboloid projection in the
// VPL's lookAtRH basis (up = (0, 1, 0.01)):
//   d = pixel - vp,  s = q / |q| with q = vn x up,  u = s x vn,
//   c = ((s.d, u.d) / |d| / (1 + vn.d / |d|) + 1) * S / 2,  texel = floor(c)
// (final_gathering.comp:118-127, ism_utils.glsl:5-40).
// (cx, cy) are estimates of c within 0.5 texel (the FP32 evaluation's, off
// by ~1e-5); each coordinate's floor is then k or k - 1 for k = rint(c_est),
// and it is k exactly when
//   S (q.d) - (2k - S) |q| (|d| + vn.d) >= 0        (x; cross(q, vn).d for y)
// since |q| (|d| + vn.d) > 0. The sign is evaluated without any division:
// d is taken homogeneously, (r.xyz - vp r.w) sign(r.w) for r = m ndc with
// ndc scaled by w h, which leaves every ratio above unchanged. The decision
// agrees with an FP64 evaluation except within ~1e-11 texel of an edge.
MLGI_FF void texel_exact(const float* m, int w, int h, int x, int y, float depth, const float* vp, const float* vn,
                         int S, float cx, float cy, int* tx, int* ty) {
  MLGI_FF_EXACT
  // w h ndc = ((2x - w) h, (2y - h) w, (2 depth - 1) w h, w h)
  const F wh = two_prod((float)w, (float)h);
  const F c0 = two_prod((float)(2 * x - w), (float)h);
  const F c1 = two_prod((float)(2 * y - h), (float)w);
  const F c2 = mul(two_sum(depth * 2.0f, -1.0f), wh);  // depth * 2 is exact
  F r[4];
  for (int row = 0; row < 4; ++row)
    r[row] = add(add(mul(c0, m[row]), mul(c1, m[4 + row])), add(mul(c2, m[8 + row]), mul(wh, m[12 + row])));
  F3 d{sub(r[0], mul(r[3], vp[0])), sub(r[1], mul(r[3], vp[1])), sub(r[2], mul(r[3], vp[2]))};
  if (r[3].hi < 0.0f) d = F3{neg(d.x), neg(d.y), neg(d.z)};
  // q = vn x up, up = (0, 1, 0.01)
  const float upz = 0.01f;
  const F3 q{add(two_prod(vn[1], upz), -vn[2]), neg(two_prod(vn[0], upz)), from(vn[0])};
  const F3 n{from(vn[0]), from(vn[1]), from(vn[2])};
  const F3 qn{sub(mul(q.y, n.z), mul(q.z, n.y)), sub(mul(q.z, n.x), mul(q.x, n.z)), sub(mul(q.x, n.y), mul(q.y, n.x))};
  const F scale = mul(sqrt(dot(q, q)), add(sqrt(dot(d, d)), dot(n, d)));  // |q| (|d| + vn.d)
  const int kx = (int)rintf(cx), ky = (int)rintf(cy);
  const F ex = sub(mul(dot(q, d), (float)S), mul(scale, (float)(2 * kx - S)));
  const F ey = sub(mul(dot(qn, d), (float)S), mul(scale, (float)(2 * ky - S)));
  *tx = ex.hi >= 0.0f ? kx : kx - 1;
  *ty = ey.hi >= 0.0f ? ky : ky - 1;
}

}  // namespace ff
}  // namespace mlgi
