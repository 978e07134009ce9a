// Interleaved final gathering with per-tile VPL culling.
// Follows gi/final_gathering.comp:63-135 (brute-force class list) and the
// dispatch of GIStage::compute_final_gathering (GIStage.cpp:261-306).
//
// Work decomposition. Pixel (x, y) belongs to interleave class
// k = (x mod I) + (y mod I) * I and sums the VPLs [k*N/I^2, (k+1)*N/I^2).
// A workgroup owns a TW x TH tile of one class's pixels (pixels I apart on
// screen, like the reference's 8x8 workgroups); each wave64 of it owns 64 of
// those pixels, so every lane of a wave walks the same VPL list, in chunks of
// 64. Each surviving VPL's record is wave-uniform and read through the
// scalar data cache into scalar registers (a broadcast LDS read returns the
// record to all 64 lanes, 1 KiB of LDS bandwidth per 16 bytes, which bound
// the gather at 4K).
//
// Culling (replaces the clustered light lists, light_lists.comp:104-119).
// Each wave reduces the bounding box of its pixels' positions; lane j tests
// VPL j of a 64-VPL chunk against it: when the VPL's tangent plane has the
// whole box behind it, by a margin far above FP32 rounding, every pixel would
// get angleFactor == 0 from that VPL. VPLs of zero flux are dropped too. A
// ballot turns the tests into a 64-bit survivor mask the wave iterates. (A
// second test, the VPL behind the tangent plane of every pixel from a box of
// the pixels' normals, cut the 1080p survivors only from 611 K to 589 K and
// cost its seven more reductions per wave: the gather was 4.6% slower with
// it, and a survivor that lights no lane only costs its hemisphere tests.)
//
// Per pair, the hemisphere tests come first and use only fused multiply-adds
// against the VPL's plane offset n.p (stored in the record) and the pixel's
// own n.p; the reciprocal square root, projection and ISM gather run only for
// lit pairs. The geometry term is written as -(n_v.d)(n_p.d)/|d|^6, and the
// paraboloid coordinate (s.d/|d|) / (1 - v.z) as (s.d) / (|d| + n_v.d), one
// reciprocal for both axes, evaluated in ISM-local texels so that no FP32
// rounding at atlas scale enters the NEAREST decision.
//
// Workgroup order is XCD-aware: consecutive workgroups go to different XCDs
// (round robin), so the linear work index is remapped such that the I^2
// classes of one screen tile, which read the same G-buffer cache lines, land
// on one XCD and its L2.
#include "batch_device.hpp"
#include "float_float.hpp"
#include "kernels.hpp"

namespace mlgi {
namespace {

constexpr int kXcds = 8;

// VPL fields the gather uses (72 bytes of the 128-byte record).
struct VplLds {
  float px, py, pz, plane;
  float nx, ny, nz, sx;
  float sy, sz, ux, uy;
  float uz, cr, cg, cb;
  uint32_t hrg, hb;  // the flux as halves: r | g << 16, b (record slots 11, 15)
  uint32_t zero;     // zero-flux flags (slot 7): bit 0 floats, bit 1 halves
};

// The bits of a float taken by value: __builtin_bit_cast applied directly to
// a vector element (v.w) reads the wrong element with this compiler.
__device__ __forceinline__ uint32_t float_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

__device__ __forceinline__ VplLds load_vpl(const float* rec) {
  const float4* r = reinterpret_cast<const float4*>(rec);
  const float4 a = gload4(r), b = gload4(r + 1), c = gload4(r + 2), d = gload4(r + 3), e = gload4(r + 4);
  return VplLds{a.x, a.y, a.z, a.w, b.x, b.y, b.z, c.x, c.y, c.z, d.x, d.y, d.z, e.x, e.y, e.z,
                float_bits(c.w), float_bits(d.w), float_bits(b.w)};
}

// Wave64 min / max, result wave-uniform. Floats are mapped to unsigned keys
// of the same order, so each step is one unsigned min/max with a DPP operand
// (no NaN canonicalisation). DPP steps: swap within pairs and quads, mirror
// within half-rows and rows (every lane then holds its row's value), then
// broadcast lane 15 into rows 1 and 3 and lane 31 into rows 2 and 3; lane 63
// ends with the whole wave's value. Lanes of rows a broadcast skips read the
// op's identity. Every lane must be active.
__device__ __forceinline__ uint32_t order_key(float f) {
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
}
__device__ __forceinline__ float from_order_key(uint32_t k) {
  return __builtin_bit_cast(float, k ^ ((uint32_t)((int32_t)~k >> 31) | 0x80000000u));
}
template <int kCtrl, int kRowMask, bool kMax>
__device__ __forceinline__ uint32_t dpp_step(uint32_t v) {
  const uint32_t identity = kMax ? 0u : 0xFFFFFFFFu;
  const uint32_t o = (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)v, kCtrl, kRowMask, 0xF, false);
  return kMax ? (v > o ? v : o) : (v < o ? v : o);
}
constexpr int kDppQuadSwap1 = 0xB1;     // quad_perm [1,0,3,2]
constexpr int kDppQuadSwap2 = 0x4E;     // quad_perm [2,3,0,1]
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140;
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;

template <bool kMax>
__device__ __forceinline__ float wave_reduce(float f) {
  uint32_t v = order_key(f);
  v = dpp_step<kDppQuadSwap1, 0xF, kMax>(v);
  v = dpp_step<kDppQuadSwap2, 0xF, kMax>(v);
  v = dpp_step<kDppRowHalfMirror, 0xF, kMax>(v);
  v = dpp_step<kDppRowMirror, 0xF, kMax>(v);
  v = dpp_step<kDppRowBcast15, 0xA, kMax>(v);
  v = dpp_step<kDppRowBcast31, 0xC, kMax>(v);
  return from_order_key((uint32_t)__builtin_amdgcn_readlane((int)v, 63));
}
__device__ __forceinline__ float wave_min(float v) { return wave_reduce<false>(v); }
__device__ __forceinline__ float wave_max(float v) { return wave_reduce<true>(v); }

struct Pixel {
  int x, y;
  bool valid;
  float depth;
  float fx, fy, fz;  // world position
  float nx, ny, nz;  // decoded face normal
  float plane;       // n_p . position
};

// Pixel (cx, cy) of class (kx, ky) and its reconstruction:
// ndc = vec4(uv, depth, 1)*2-1, world = VPinv*ndc / w (depth column first).
__device__ __forceinline__ Pixel load_pixel(const GatherArgs& a, int cx, int cy, int kx, int ky) {
  Pixel p;
  p.x = cx * a.interleave + kx;
  p.y = cy * a.interleave + ky;
  p.valid = p.x < a.width && p.y < a.height;
  const int sx = p.valid ? p.x : 0, sy = p.valid ? p.y : 0;
  const size_t idx = (size_t)sy * a.width + sx;
  const float d = gload(a.depth + idx);
  p.depth = d;
  const float n0 = gload(a.normal + idx * 3), n1 = gload(a.normal + idx * 3 + 1), n2 = gload(a.normal + idx * 3 + 2);
  // The texture coordinate and the perspective divide use reciprocals (the
  // screen size's once per wave): the position moves by a few ulps against
  // the divisions, far inside the gather's tolerance. The exact texel
  // decision of the FP32 form rebuilds the position from (x, y, depth)
  // itself.
  const float rw = __builtin_amdgcn_rcpf((float)a.width), rh = __builtin_amdgcn_rcpf((float)a.height);
  const float u = (float)sx * rw, v = (float)sy * rh;
  const float c0 = u * 2.0f - 1.0f, c1 = v * 2.0f - 1.0f, c2 = d * 2.0f - 1.0f;
  const float* m = a.vp_inv.m;
  float r[4];
#pragma unroll
  for (int row = 0; row < 4; ++row)
    r[row] = __builtin_fmaf(m[row], c0,
                            __builtin_fmaf(m[4 + row], c1, __builtin_fmaf(m[8 + row], c2, m[12 + row])));
  const float rw3 = __builtin_amdgcn_rcpf(r[3]);
  p.fx = r[0] * rw3;
  p.fy = r[1] * rw3;
  p.fz = r[2] * rw3;
  p.nx = n0 * 2.0f - 1.0f;
  p.ny = n1 * 2.0f - 1.0f;
  p.nz = n2 * 2.0f - 1.0f;
  p.plane = __builtin_fmaf(p.nx, p.fx, __builtin_fmaf(p.ny, p.fy, p.nz * p.fz));
  return p;
}

struct WaveBounds {
  float cx, cy, cz, hx, hy, hz;  // position box centre / half extent
  float scale;                   // coordinate magnitude for the margins
};

// Position box of the wave's valid pixels; false when the wave has no pixel
// on screen. Wave-uniform.
__device__ __forceinline__ bool wave_bounds(const Pixel& p, WaveBounds& wb) {
  const float inf = __builtin_inff();
  const float mnx = wave_min(p.valid ? p.fx : inf), mxx = wave_max(p.valid ? p.fx : -inf);
  const float mny = wave_min(p.valid ? p.fy : inf), mxy = wave_max(p.valid ? p.fy : -inf);
  const float mnz = wave_min(p.valid ? p.fz : inf), mxz = wave_max(p.valid ? p.fz : -inf);
  wb.cx = 0.5f * (mnx + mxx);
  wb.cy = 0.5f * (mny + mxy);
  wb.cz = 0.5f * (mnz + mxz);
  wb.hx = 0.5f * (mxx - mnx) * 1.0001f;
  wb.hy = 0.5f * (mxy - mny) * 1.0001f;
  wb.hz = 0.5f * (mxz - mnz) * 1.0001f;
  wb.scale = 4.0f * fmaxf(fmaxf(fmaxf(fabsf(mnx), fabsf(mxx)), fmaxf(fabsf(mny), fabsf(mxy))),
                          fmaxf(fabsf(mnz), fabsf(mxz)));
  return mnx <= mxx;
}

// Culling test of one VPL against a wave's bounds: false when every pixel of
// the wave gets angleFactor == 0 from it, or when its flux is zero (an RSM
// texel outside the spot cone; MIXED: zero after rounding to halves, the
// values the sums multiply): every contribution is then flux * weight = +0
// with a finite weight, which leaves the sums' bits unchanged. Reads only
// the record's first 32 bytes (the zero-flux flags sit beside the normal):
// every load per lane is paid by every wave for every chunk.
template <bool kMixed>
__device__ __forceinline__ bool cull_keep(const VplLds& v, const WaveBounds& w) {
  const bool dark = (v.zero >> (kMixed ? 1 : 0)) & 1u;
  // max over the box of n . (p - p_vpl) = n . (c - p_vpl) + |n| . h
  const float dx = w.cx - v.px, dy = w.cy - v.py, dz = w.cz - v.pz;
  const float ax = fabsf(v.nx), ay = fabsf(v.ny), az = fabsf(v.nz);
  const float spread = ax * w.hx + ay * w.hy + az * w.hz;
  const float support = v.nx * dx + v.ny * dy + v.nz * dz + spread;
  const float mag = ax * fabsf(dx) + ay * fabsf(dy) + az * fabsf(dz) + spread;
  const float tol = 1e-4f * mag + 1e-6f * (w.scale + fabsf(v.plane));
  const bool behind_vpl = support < -tol;
  return !(behind_vpl || dark);
}

// Survivor mask of the `count` VPLs of a chunk, lane j testing VPL j, from
// the chunk's records in global memory (flux as floats). With culling off
// (a wave-uniform flag) every VPL of the chunk survives.
template <bool kMixed>
__device__ __forceinline__ unsigned long long cull_chunk(const float* crec, int count, const WaveBounds& w,
                                                         int cull) {
  if (!cull) return count >= 64 ? ~0ull : ((1ull << count) - 1ull);
  const int lane = threadIdx.x & 63;
  bool keep = false;
  if (lane < count) keep = cull_keep<kMixed>(load_vpl(crec + (size_t)lane * kVplFloats), w);
  return __ballot(keep);
}

// NEAREST texel of the pair re-decided exactly: the FP32 texel coordinate
// lies within kTie of a texel edge, closer than its rounding error can be
// trusted. The FP32 estimate names the nearest edge; which side of it the
// coordinate lies on is decided in float-float arithmetic (FP32 operations
// only, ~44 bits; float_float.hpp) from the FP32 inputs, as an FP64
// evaluation would decide except within ~1e-11 texel of the edge.
constexpr float kTie = 2.5e-4f;

// Out of line: the float-float temporaries would otherwise raise the register
// count of the whole kernel for a path taken by well under 1% of pairs.
// Returns the texel as (ty << 16) | tx.
__device__ __noinline__ int refine_texel(Mat4 m4, int width, int height, int ism_side, int x, int y,
                                         float depth, float vpx, float vpy, float vpz, float vnx,
                                         float vny, float vnz, float cx, float cy) {
  const float vp[3] = {vpx, vpy, vpz}, vn[3] = {vnx, vny, vnz};
  int tx, ty;
  ff::texel_exact(m4.m, width, height, x, y, depth, vp, vn, ism_side, cx, cy, &tx, &ty);
  return ((ty & 0xFFFF) << 16) | (tx & 0xFFFF);
}

struct VplTest {
  float px, py, pz, plane, nx, ny, nz;
};

struct LitConsts {
  __amdgpu_buffer_rsrc_t ism;  // the unorm16 ISM atlas as a raw buffer
  float hs;   // ((p + 1) / 2) * S = p * S/2 + S/2
  int amask;  // atlas side - 1
  int alog;   // log2(atlas side)
  float k1;   // smoothstep slope / zFar
  float k2;   // smoothstep slope / 65535
};

// The ISM is fetched through a raw buffer resource (base and range in scalar
// registers, a 32-bit byte offset per lane; texel < 2^28), which saves the
// 64-bit address arithmetic of a global load per lit pair. Resource word 3
// 0x00020000 selects the 32-bit data format raw buffer loads use on CDNA.
__device__ __forceinline__ LitConsts lit_consts(const GatherArgs& a) {
  return LitConsts{__builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t*>(a.ism), 0, 0x7FFFFFFF, 0x00020000),
                   0.5f * (float)a.ism_side, a.atlas_side - 1, 31 - __builtin_clz(a.atlas_side),
                   a.shadow_k1, a.shadow_k2};
}

// Atlas texel of ISM texel (tx, ty) in the tile at (ox, oy), wrapped as
// GL_REPEAT wraps the atlas.
__device__ __forceinline__ uint32_t atlas_texel(const LitConsts& lc, int ox, int oy, int tx, int ty) {
  return (uint32_t)((((oy + ty) & lc.amask) << lc.alog) | ((ox + tx) & lc.amask));
}

// The smoothstep argument of a lit pair at ISM texel (tx, ty), clamped to
// [0, 1]: the pair's shadow weight is a function of this value alone.
__device__ __forceinline__ float shadow_arg(const LitConsts& lc, int ox, int oy, int tx, int ty, float sbase) {
  const uint16_t occ = __builtin_amdgcn_raw_buffer_load_b16(lc.ism, atlas_texel(lc, ox, oy, tx, ty) << 1, 0, 0);
  return __builtin_amdgcn_fmed3f(__builtin_fmaf((float)occ, lc.k2, sbase), 0.0f, 1.0f);
}

// One lit pair (n_v . d > 0, n_p . d < 0) up to its ISM fetch: the geometry
// term with its clamp (final_gathering.comp:100-117), the texel of the
// pair's ISM (:118-127) and the distance part of the smoothstep argument.
// (px, py, pz) is the pixel's position; x, y, depth identify the pixel for
// the exact texel decision.
struct PairSetup {
  float g;         // clamped geometry term
  float sbase;     // 1 - dist / (bias zFar): the smoothstep argument without occ
  uint32_t texel;  // ISM atlas texel (0 when shadowing is off)
};

template <bool kMixed>
__device__ __forceinline__ PairSetup pair_setup(const GatherArgs& a, float px, float py, float pz, int x, int y,
                                                float depth, const VplTest& vt, const VplLds& v, int vi, float ndot,
                                                float pdot, bool shadow, const LitConsts& lc) {
  const float dx = px - vt.px, dy = py - vt.py, dz = pz - vt.pz;
  const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
  const float inv = __builtin_amdgcn_rsqf(d2);
  const float inv2 = inv * inv;
  PairSetup r{0.0f, 0.0f, 0u};
  // cos_v cos_p / d^4 = -(ndot pdot) / d^6
  r.g = fminf(-(ndot * pdot) * inv2 * (inv2 * inv2), a.clamp_value);
  if (shadow) {
    const float dist = d2 * inv;
    const float rh = lc.hs * __builtin_amdgcn_rcpf(dist + ndot);
    const float sd = __builtin_fmaf(v.sx, dx, __builtin_fmaf(v.sy, dy, v.sz * dz));
    const float ud = __builtin_fmaf(v.ux, dx, __builtin_fmaf(v.uy, dy, v.uz * dz));
    const float cx = __builtin_fmaf(sd, rh, lc.hs);
    const float cy = __builtin_fmaf(ud, rh, lc.hs);
    // the VPL's ISM tile: ISM index vi - offset (the VPL range start with
    // scaleISMs), tile row int(index / n1d) truncated toward zero; a negative
    // index gives tiles left of or above the atlas, which the texel wrap maps
    // back as GL_REPEAT does (final_gathering.comp:43-45, ism_utils.glsl:33-37)
    const int ism = vi - a.ism_offset;
    const int trow = (ism < 0 ? ism + a.n1d - 1 : ism) >> a.n1d_log2;
    const int ox = (ism - trow * a.n1d) * a.ism_side;
    const int oy = trow * a.ism_side;
    r.sbase = __builtin_fmaf(-dist, lc.k1, 1.0f);
    int tx, ty;
    if (kMixed) {  // NEAREST decided in FP32 alone
      tx = (int)floorf(cx);
      ty = (int)floorf(cy);
    } else {
      const float fx = __builtin_amdgcn_fractf(cx), fy = __builtin_amdgcn_fractf(cy);
      tx = (int)(cx - fx);
      ty = (int)(cy - fy);
      const bool near_x = fabsf(fx - 0.5f) > 0.5f - kTie, near_y = fabsf(fy - 0.5f) > 0.5f - kTie;
      if (near_x || near_y) {
        // The exact texel is one of the (up to) four on either side of the
        // near edges. When all of them give the same clamped smoothstep
        // argument the pair's weight does not depend on the choice.
        const int ax = near_x ? (fx < 0.5f ? -1 : 1) : 0, ay = near_y ? (fy < 0.5f ? -1 : 1) : 0;
        const float t0 = shadow_arg(lc, ox, oy, tx, ty, r.sbase);
        const bool same = shadow_arg(lc, ox, oy, tx + ax, ty, r.sbase) == t0 &&
                          shadow_arg(lc, ox, oy, tx, ty + ay, r.sbase) == t0 &&
                          shadow_arg(lc, ox, oy, tx + ax, ty + ay, r.sbase) == t0;
        if (!same) {
          const int t2 = refine_texel(a.vp_inv, a.width, a.height, a.ism_side, x, y, depth, vt.px, vt.py, vt.pz,
                                      vt.nx, vt.ny, vt.nz, cx, cy);
          tx = (int)(int16_t)(t2 & 0xFFFF);
          ty = t2 >> 16;
        }
      }
    }
    r.texel = atlas_texel(lc, ox, oy, tx, ty);
  }
  return r;
}

// The weight after the ISM fetch: smoothstep(1 - bias, 1, 1 - (dist/zFar -
// occ)) with t = 1 + (occ - dist/zFar) / bias (final_gathering.comp:128-131).
__device__ __forceinline__ float pair_finish(const PairSetup& s, uint16_t occ16, const LitConsts& lc) {
  float t = __builtin_fmaf((float)occ16, lc.k2, s.sbase);
  t = __builtin_amdgcn_fmed3f(t, 0.0f, 1.0f);  // shadow_arg's value
  return s.g * (t * t * __builtin_fmaf(-2.0f, t, 3.0f));
}

// Weight of one lit pair, fetch included.
template <bool kMixed>
__device__ __forceinline__ float pair_weight(const GatherArgs& a, float px, float py, float pz, int x, int y,
                                             float depth, const VplTest& vt, const VplLds& v, int vi, float ndot,
                                             float pdot, bool shadow, const LitConsts& lc) {
  const PairSetup s = pair_setup<kMixed>(a, px, py, pz, x, y, depth, vt, v, vi, ndot, pdot, shadow, lc);
  if (!shadow) return s.g;
  return pair_finish(s, __builtin_amdgcn_raw_buffer_load_b16(lc.ism, s.texel << 1, 0, 0), lc);
}

// The flux half h of a record (binary16 bits) as a float: the conversion
// folds into the mixed-precision FMA that multiplies it (v_fma_mix_f32).
__device__ __forceinline__ float half_bits(uint32_t h) { return (float)__builtin_bit_cast(_Float16, (uint16_t)h); }

// Contribution of one lit pair to the pixel's running sums (one fused
// multiply-add per channel; MIXED: the flux as the record's halves).
template <bool kMixed>
__device__ __forceinline__ void accumulate_pair(const GatherArgs& a, const Pixel& p, const VplTest& vt,
                                                const VplLds& v, int vi, float ndot, float pdot,
                                                bool shadow, const LitConsts& lc, float& ar, float& ag,
                                                float& ab) {
  const float g = pair_weight<kMixed>(a, p.fx, p.fy, p.fz, p.x, p.y, p.depth, vt, v, vi, ndot, pdot, shadow, lc);
  const float fr = kMixed ? half_bits(v.hrg & 0xFFFFu) : v.cr;
  const float fg = kMixed ? half_bits(v.hrg >> 16) : v.cg;
  const float fb = kMixed ? half_bits(v.hb & 0xFFFFu) : v.cb;
  ar = __builtin_fmaf(fr, g, ar);
  ag = __builtin_fmaf(fg, g, ag);
  ab = __builtin_fmaf(fb, g, ab);
}

// Record of VPL vi (a wave-uniform index) read through the scalar data
// cache into scalar registers. No LDS traffic, and the VALU takes the fields
// as scalar operands.
__device__ __forceinline__ VplLds uniform_vpl(const float* vpl, int vi) {
  typedef float cf4 __attribute__((ext_vector_type(4)));
  typedef const __attribute__((address_space(4))) cf4 ConstF4;
  const int u = __builtin_amdgcn_readfirstlane(vi);
  ConstF4* r = (ConstF4*)(vpl + (size_t)u * kVplFloats);
  const cf4 a = r[0], b = r[1], c = r[2], d = r[3], e = r[4];
  return VplLds{a.x, a.y, a.z, a.w, b.x, b.y, b.z, c.x, c.y, c.z, d.x, d.y, d.z, e.x, e.y, e.z,
                float_bits(c.w), float_bits(d.w), float_bits(b.w)};
}

// The hemisphere-test fields of VPL vi (wave-uniform), as uniform_vpl.
__device__ __forceinline__ VplTest uniform_test(const float* vpl, int vi) {
  typedef float cf4 __attribute__((ext_vector_type(4)));
  typedef const __attribute__((address_space(4))) cf4 ConstF4;
  const int u = __builtin_amdgcn_readfirstlane(vi);
  ConstF4* r = (ConstF4*)(vpl + (size_t)u * kVplFloats);
  const cf4 a = r[0], b = r[1];
  return VplTest{a.x, a.y, a.z, a.w, b.x, b.y, b.z};
}

// Kahan step: total += x.
__device__ __forceinline__ void kahan_add(float& total, float& comp, float x) {
  const float y = x - comp;
  const float t = total + y;
  comp = (t - total) - y;
  total = t;
}

// Record fetch variants of the single-frame kernel (picked per shape by the
// autotuner); both read the wave-uniform record through scalar loads:
//   kFetchWhole     the full record at the top of each iteration
//   kFetchTestFirst the hemisphere-test fields per iteration, the rest only
//                   when a lane is lit
enum FetchVariant { kFetchWhole = 0, kFetchTestFirst = 1 };

// Precision forms.
//   FP32  the reference's arithmetic in FP32, NEAREST texel decisions that
//         FP32 cannot resolve re-decided in float-float arithmetic
//         (rel-L2 <= 1e-4; no FP64 anywhere).
//   MIXED the flux as halves (rounded once, at VPL extraction) multiplied by
//         f16 x f32 + f32 mixed FMAs, NEAREST decided in FP32 alone,
//         per-pixel sums in FP32 per 64-VPL chunk, chunks combined with
//         Kahan compensation (rel-L2 <= 2e-3). The geometry, projection and
//         distance terms stay FP32: an FP16 basis moves texel decisions by up
//         to 0.03 texel, which alone costs ~6e-3 rel-L2, and packed-half
//         geometry and smoothstep terms on pairs of VPLs measured slower
//         (more scalar work pairing the VPLs than VALU saved) and 2.5x less
//         precise.

// Sums of one pixel over its class's VPLs [first, first + per_class), in
// chunks of 64: the culling survivors of each chunk in order, the hemisphere
// tests first, the rest of the pair only where a lane is lit. MIXED: each
// chunk's FP32 sums are added to the totals with Kahan compensation.
template <bool kMixed, bool kShadow, int kFetch>
__device__ __forceinline__ void class_sums(const GatherArgs& a, const Pixel& p, const WaveBounds& wb, int first,
                                           const LitConsts& lc, float& sr, float& sg, float& sb) {
  float ar = 0.0f, ag = 0.0f, ab = 0.0f;  // FP32: the sums; MIXED: this chunk's sums
  float tr = 0.0f, tg = 0.0f, tb = 0.0f;  // MIXED: Kahan totals
  float er = 0.0f, eg = 0.0f, eb = 0.0f;  // MIXED: Kahan compensations
  for (int chunk = 0; chunk < a.per_class; chunk += 64) {
    const int count = min(64, a.per_class - chunk);
    const float* crec = a.vpl + (size_t)(first + chunk) * kVplFloats;
    unsigned long long mask = cull_chunk<kMixed>(crec, count, wb, a.cull);
    while (mask) {
      const int j = __builtin_ctzll(mask);
      mask &= mask - 1ull;
      const int vi = first + chunk + j;
      VplTest vt;
      VplLds vf;
      if (kFetch == kFetchWhole) {
        vf = uniform_vpl(a.vpl, vi);
        vt = VplTest{vf.px, vf.py, vf.pz, vf.plane, vf.nx, vf.ny, vf.nz};
      } else {
        vt = uniform_test(a.vpl, vi);
      }
      // n_vpl . (p - p_vpl) and n_p . (p - p_vpl) from the plane offsets
      const float ndot = __builtin_fmaf(vt.nx, p.fx, __builtin_fmaf(vt.ny, p.fy, __builtin_fmaf(vt.nz, p.fz, -vt.plane)));
      const float pdot = __builtin_fmaf(-p.nx, vt.px, __builtin_fmaf(-p.ny, vt.py, __builtin_fmaf(-p.nz, vt.pz, p.plane)));
      if (ndot > 0.0f && pdot < 0.0f) {
        const VplLds v = (kFetch == kFetchTestFirst) ? uniform_vpl(a.vpl, vi) : vf;
        accumulate_pair<kMixed>(a, p, vt, v, vi, ndot, pdot, kShadow, lc, ar, ag, ab);
      }
    }
    if (kMixed) {
      kahan_add(tr, er, ar);
      kahan_add(tg, eg, ag);
      kahan_add(tb, eb, ab);
      ar = ag = ab = 0.0f;
    }
  }
  sr = kMixed ? tr : ar;
  sg = kMixed ? tg : ag;
  sb = kMixed ? tb : ab;
}

template <int TW, int TH, bool kShadow, int kFetch, bool kMixed>
__global__ __launch_bounds__(TW* TH) void final_gather_kernel(GatherArgs a, int tiles_x,
                                                              int num_work, int per_xcd) {
  const int b = blockIdx.x;
  const int w = (b % kXcds) * per_xcd + b / kXcds;
  if (w >= num_work) return;
  const int classes_log2 = 2 * a.interleave_log2;
  const int k = w & ((1 << classes_log2) - 1);
  const int tile = w >> classes_log2;
  const int kx = k & (a.interleave - 1), ky = k >> a.interleave_log2;
  const int tile_x = tile % tiles_x, tile_y = tile / tiles_x;
  const int lx = threadIdx.x % TW, ly = threadIdx.x / TW;
  const Pixel p = load_pixel(a, tile_x * TW + lx, tile_y * TH + ly, kx, ky);

  // ---- wave bounds for culling
  WaveBounds wb;
  const bool wave_live = wave_bounds(p, wb);  // some pixel of this wave is on screen (wave-uniform)

  const int first = k * a.per_class;
  const LitConsts lc = lit_consts(a);
  float ar = 0.0f, ag = 0.0f, ab = 0.0f;
  if (wave_live) class_sums<kMixed, kShadow, kFetch>(a, p, wb, first, lc, ar, ag, ab);
  if (p.valid) {
    float* o = a.out + ((size_t)p.y * a.width + p.x) * 3;
    gstore(o, ar * a.out_scale);
    gstore(o + 1, ag * a.out_scale);
    gstore(o + 2, ab * a.out_scale);
  }
}

// Batched gather: one wave64 workgroup per (frame, class, strip) item, over
// all frames of one precision in one launch. A strip is 2^w x 64/2^w pixels of one
// interleave class (w per frame, picked by the autotuner). Waves share no
// memory and never synchronise, so frames of different shapes share
// workgroups freely. The per-pixel arithmetic and summation order are the
// single-frame kernel's, so a frame gives the same bits alone or in any
// batch.
//
// (Compacting the lit pairs of a wave into an LDS queue, so that lanes spend
// the projection and ISM fetch on lit pairs only, measured 4.4x slower at
// 1080p; each queued pair moved ~130 bytes through LDS. Issuing the ISM
// fetches of four VPLs back to back measured 3% slower for the mixed form,
// and 47% slower for FP32, whose FP64 tie path then spills. A two-stage
// software pipeline, the next pair's setup and fetch issued before the
// previous pair is finished with two alternating register sets, measured 7%
// slower at 1080p even with survivors that light no lane filtered out first
// (43% slower without the filter), and equal at 4K: with eight waves per
// SIMD the fetch latency is already hidden, and the pipeline adds work.
// Letting each lane walk its own lit VPLs (a per-lane 64-bit lit mask from
// the hemisphere tests, then per-lane record reads from 17-word LDS slots),
// so that a wave runs as many lit passes as its busiest lane, measured 9%
// slower at 1080p and 7% slower at 4K: per-lane record reads and tile
// addressing cost more than the passes it saves. Requesting the first
// chunk's culling records before the pixel loads, to overlap the two
// latencies, measured 9% slower at 4K and 15% slower for FP32: the records
// stay live across the pixel reconstruction and the bound reductions.)
// Occupancy: eight waves per SIMD (64 VGPRs) for MIXED; the FP32 form's
// exact-texel path needs 67, which spill at eight.
template <bool kMixed>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kMixed ? 8 : 7))) void batch_gather_kernel(const BatchFrame* __restrict__ frames, int n,
                                                           uint32_t total, int chunk_log2) {
  const int lane = threadIdx.x;
  // One single-wave workgroup per item, dispatched by the hardware as waves
  // retire (a persistent grid walking the items ran the same instructions
  // 20% slower). Items are dealt to the XCDs in chunks of 2^chunk_log2
  // consecutive items (bands of neighbouring strips with all their
  // interleave classes, which share G-buffer lines), chunk c to XCD c mod 8:
  // workgroup b runs on XCD b mod 8 and takes that XCD's (b / 8)-th item.
  // Every XCD works in several bands across the screen, so strips of very
  // different cost (lit or not) balance over the chip.
  const uint32_t xcd = blockIdx.x % kBatchXcds, i = blockIdx.x / kBatchXcds;
  const uint32_t chunk_mask = (1u << chunk_log2) - 1u;
  const uint32_t item = (((i >> chunk_log2) * kBatchXcds + xcd) << chunk_log2) | (i & chunk_mask);
  if (item >= total) return;  // the tail of the last chunks
  {
    const BatchFrame& f = frames[batch_find<&BatchFrame::first_wave>(frames, n, item)];
    const GatherArgs a = f.g;  // by value: the fields live in registers, not re-read per VPL
    const uint32_t local = item - f.first_wave;
    const int classes_log2 = 2 * a.interleave_log2;
    const int k = local & ((1u << classes_log2) - 1);
    const int strip = local >> classes_log2;
    const int kx = k & (a.interleave - 1), ky = k >> a.interleave_log2;
    const int sx = strip % f.strips_x, sy = strip / f.strips_x;
    const int wlog = f.wave_w_log2;
    const int lx = lane & ((1 << wlog) - 1), ly = lane >> wlog;
    const Pixel p = load_pixel(a, (sx << wlog) + lx, sy * (64 >> wlog) + ly, kx, ky);
    WaveBounds wb;
    const bool live = wave_bounds(p, wb);
    if (!live) return;  // wave-uniform: nothing of this strip is on screen
    const int first = k * a.per_class;
    const LitConsts lc = lit_consts(a);
    float ar = 0.0f, ag = 0.0f, ab = 0.0f;
    // the frame's shadowing flag picks one of two instances of the VPL loop
    // (a scalar branch), so that the lit path holds no per-lane copy of it;
    // the loop is the single-frame kernel's with the whole-record fetch
    if (__builtin_amdgcn_readfirstlane(a.shadowing))
      class_sums<kMixed, true, kFetchWhole>(a, p, wb, first, lc, ar, ag, ab);
    else
      class_sums<kMixed, false, kFetchWhole>(a, p, wb, first, lc, ar, ag, ab);
    if (p.valid) {
      float* o = a.out + ((size_t)p.y * a.width + p.x) * 3;
      gstore(o, ar * a.out_scale);
      gstore(o + 1, ag * a.out_scale);
      gstore(o + 2, ab * a.out_scale);
    }
  }
}

// Lit-pair counter: pairs (pixel, VPL of its class) with angleFactor > 0,
// evaluated directly as max(0, n_v.d) * max(0, -n_p.d) > 0.
__global__ __launch_bounds__(64) void count_lit_kernel(GatherArgs a, int tiles_x, int num_work,
                                                       unsigned long long* counter) {
  const int w = blockIdx.x;
  if (w >= num_work) return;
  const int classes_log2 = 2 * a.interleave_log2;
  const int k = w & ((1 << classes_log2) - 1);
  const int tile = w >> classes_log2;
  const int kx = k & (a.interleave - 1), ky = k >> a.interleave_log2;
  const int tile_x = tile % tiles_x, tile_y = tile / tiles_x;
  const Pixel p = load_pixel(a, tile_x * 8 + (threadIdx.x & 7), tile_y * 8 + (threadIdx.x >> 3), kx, ky);
  unsigned int lit = 0;
  const int first = k * a.per_class;
  for (int j = 0; j < a.per_class; ++j) {
    const VplLds v = load_vpl(a.vpl + (size_t)(first + j) * kVplFloats);
    const float dx = p.fx - v.px, dy = p.fy - v.py, dz = p.fz - v.pz;
    const float ndot = v.nx * dx + v.ny * dy + v.nz * dz;
    const float pdot = p.nx * dx + p.ny * dy + p.nz * dz;
    lit += (p.valid && ndot > 0.0f && pdot < 0.0f) ? 1u : 0u;
  }
  unsigned long long s = lit;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0)
    __hip_atomic_fetch_add(counter, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int TW, int TH, int kFetch, bool kMixed>
hipError_t launch_tiled(hipStream_t s, const GatherArgs& a) {
  const int cw = (a.width + a.interleave - 1) / a.interleave;
  const int ch = (a.height + a.interleave - 1) / a.interleave;
  const int tiles_x = (cw + TW - 1) / TW, tiles_y = (ch + TH - 1) / TH;
  const int num_work = tiles_x * tiles_y * a.interleave * a.interleave;
  const int per_xcd = (num_work + kXcds - 1) / kXcds;
  const dim3 grid(per_xcd * kXcds), block(TW * TH);
  if (a.shadowing)
    hipLaunchKernelGGL((final_gather_kernel<TW, TH, true, kFetch, kMixed>), grid, block, 0, s, a,
                       tiles_x, num_work, per_xcd);
  else
    hipLaunchKernelGGL((final_gather_kernel<TW, TH, false, kFetch, kMixed>), grid, block, 0, s, a,
                       tiles_x, num_work, per_xcd);
  return hipGetLastError();
}

template <int TW, int TH>
hipError_t launch_variant(hipStream_t s, const GatherArgs& a, int fetch, bool mixed) {
  switch (fetch * 2 + (mixed ? 1 : 0)) {
    case kFetchWhole * 2: return launch_tiled<TW, TH, kFetchWhole, false>(s, a);
    case kFetchWhole * 2 + 1: return launch_tiled<TW, TH, kFetchWhole, true>(s, a);
    case kFetchTestFirst * 2: return launch_tiled<TW, TH, kFetchTestFirst, false>(s, a);
    case kFetchTestFirst * 2 + 1: return launch_tiled<TW, TH, kFetchTestFirst, true>(s, a);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_final_gather(hipStream_t s, const GatherArgs& a, int precision, int tile_w,
                               int tile_h, int fetch) {
  const bool mixed = precision == 1;
  if (tile_w == 8 && tile_h == 8) return launch_variant<8, 8>(s, a, fetch, mixed);
  if (tile_w == 16 && tile_h == 16) return launch_variant<16, 16>(s, a, fetch, mixed);
  if (tile_w == 32 && tile_h == 8) return launch_variant<32, 8>(s, a, fetch, mixed);
  if (tile_w == 16 && tile_h == 4) return launch_variant<16, 4>(s, a, fetch, mixed);
  return hipErrorInvalidValue;
}

hipError_t launch_batch_gather(hipStream_t s, const BatchFrame* frames, int n, uint32_t total,
                               int precision) {
  if (total == 0) return hipSuccess;
  // Chunks of 1024 items, halved until there are at least two per XCD: a
  // 256^2 frame (1,024 strips) in one 1024-item chunk would run on one XCD.
  // Measured at 1080p / 4K (32,400 / 129,600 items, so 1024 there): 1024-item
  // chunks 66 / 345 us, 64-item chunks 74 / 346 us, one contiguous eighth per
  // XCD 71 / 359 us.
  int chunk_log2 = 10;
  while (chunk_log2 > 0 && total < 2u * kBatchXcds * (1u << chunk_log2)) --chunk_log2;
  // every item's workgroup: whole chunks on every XCD
  const uint64_t per_round = (uint64_t)kBatchXcds << chunk_log2;
  const unsigned grid = (unsigned)((total + per_round - 1) / per_round * per_round);
  if (precision == 1)
    hipLaunchKernelGGL(batch_gather_kernel<true>, dim3(grid), dim3(64), 0, s, frames, n, total, chunk_log2);
  else
    hipLaunchKernelGGL(batch_gather_kernel<false>, dim3(grid), dim3(64), 0, s, frames, n, total, chunk_log2);
  return hipGetLastError();
}

hipError_t launch_count_lit_pairs(hipStream_t s, const GatherArgs& a,
                                  unsigned long long* counter) {
  const int cw = (a.width + a.interleave - 1) / a.interleave;
  const int ch = (a.height + a.interleave - 1) / a.interleave;
  const int tiles_x = (cw + 7) / 8, tiles_y = (ch + 7) / 8;
  const int num_work = tiles_x * tiles_y * a.interleave * a.interleave;
  hipLaunchKernelGGL(count_lit_kernel, dim3(num_work), dim3(64), 0, s, a, tiles_x, num_work, counter);
  return hipGetLastError();
}

}  // namespace mlgi
