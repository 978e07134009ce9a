// Pull-push hole filling of the ISM atlas, fused per ISM.
// Follows ism/pull.comp:20-126, ism/push.comp:44-165 and the level loop of
// ImperfectShadowmap::pullpush (ImperfectShadowmap.cpp:121-182).
// Compiled with -ffp-contract=off: the R16 result is bit-exact against the
// FP32 oracle twin.
//
// The reference runs 2L full-atlas dispatches through RGBA32F mip chains.
// Here one 256-lane workgroup owns one S x S ISM: the first pull level reads
// the u32 atlas tile (its 2x2 blocks with 8-byte loads), every pull level
// lives in LDS, the push goes back down in place (push level l overwrites
// pull level l pixel by pixel), and the level-0 push reads each pixel's
// original sample from the atlas (an L2 hit by then) and writes the unorm16
// result. 2x2 pull blocks never cross ISMs and the push border test gives
// zero weight to every texel of another ISM, so ISMs are independent and
// nothing outside the tile is read. Zero-weight terms contribute +-0 to the
// push sums, so skipping them is exact.
//
// The LDS holds the pull levels and a 1 KB table of max-depth offsets (22.9
// KB at S = 64, L = 4), so seven workgroups fit a CU: the small levels leave
// most of a workgroup's waves waiting at barriers, and other workgroups fill
// the issue slots meanwhile. (With the atlas tile and a second radius-byte
// table staged in LDS as well, 40 KB, four workgroups fit, and the 1080p
// pull-push took 102 us instead of 83.)
//
// Levels are compile-time (one kernel per ISM side and level count), so every
// level's extents, scale and the fine-to-coarse offsets are constants.
//
// Most of an atlas is empty (0xFFFFFFFF), and an empty texel decodes to a
// fixed, valid far sample: the first pull level of an all-empty 2x2 block is
// one constant E_1 (computed on the host by the same code and passed as an
// argument), written for every such block (the two-pass level 1 below).
//
// Push runs over the reference's invocations: the (up to) four fine pixels
// that share the same four coarse neighbours (push.comp:46-67), so loads,
// half unpacking and per-neighbour constants are formed once per quad. The
// quads wrap around the ISM, so that the half-empty border invocations of
// opposite sides share one quad and every quad holds four pixels (push_quad).
// The coarse-to-fine offset (coarse + 0.5) * 2 - (fine + 0.5) of each pair is
// one of +-0.5, +-1.5, exact in the reference's global coordinates as well.
// The radius tests compare a correctly rounded square root with a float,
// sqrt_rn(x) > R. Because sqrt(x) of a float x can never equal the midpoint m
// between R and the next float (m has 25 significant bits, m^2 then has more
// than 48, x has 24), that is exactly x > m^2, and for a float x exactly
// x > RD(m^2), which rd_bound forms in FP32: the same decision as the FP32
// twin without a square root, and no FP64 anywhere.
//
// Most texels of an atlas are empty: at 1080p 77% of the 2x2 blocks the
// first pull level reads are all empty, and such a block's result is the
// constant E_1. That level therefore runs in two passes: a cheap pass writes
// E_1 for the empty blocks and appends the others to a list in LDS (ballot +
// mbcnt), and the full arithmetic runs over the list, every wave iteration
// full of real work. The list lives in the LDS of the coarser levels, which
// are not written yet, so the workgroup's LDS does not grow. (The push
// levels gain nothing from the same split: once level 1 is pushed, its
// filled holes reach 88-90% of the finer quads, which then need the full
// arithmetic anyway.)
#include "batch_device.hpp"
#include "kernels.hpp"
#include "mlgi_device.hpp"

#include <cmath>

namespace mlgi {
namespace {

struct __align__(16) Texel {
  float depth, max_depth, radius;
  uint32_t disp;  // packHalf2x16 displacement
};

constexpr int kThreads = 256;

struct PullConsts {
  Texel e1;  // pull level 1 of an all-empty 2x2 block
};


__host__ __device__ inline uint32_t f2u(float f) {
  union {
    float f;
    uint32_t u;
  } x;
  x.f = f;
  return x.u;
}

// packHalf2x16 (round to nearest even). The device uses the hardware
// conversion; the host (constants of empty regions) the textbook bit method.
__host__ __device__ inline uint16_t f32_to_f16_rne(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  // The empty asm pins f as a rounded float: without it the backend folds a
  // producing fma and this conversion into one fma_mix with a single f16
  // rounding, which is not packHalf2x16 of the float result.
  float v = f;
  asm volatile("" : "+v"(v));
  const _Float16 h = (_Float16)v;
  return __builtin_bit_cast(uint16_t, h);
#else
  const uint32_t x = f2u(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7FFFFFFFu;
  if (ax >= 0x7F800000u) return (uint16_t)(sign | 0x7C00u | (ax > 0x7F800000u ? 0x200u : 0u));
  if (ax >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);
  if (ax < 0x38800000u) {
    if (ax < 0x33000000u) return (uint16_t)sign;
    const uint32_t e = ax >> 23, m = (ax & 0x7FFFFFu) | 0x800000u, sh = 126u - e;
    uint32_t q = m >> sh;
    const uint32_t rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return (uint16_t)(sign | q);
  }
  uint32_t q = (((ax >> 23) - 112u) << 10) | ((ax & 0x7FFFFFu) >> 13);
  const uint32_t rem = ax & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;
  return (uint16_t)(sign | q);
#endif
}

__host__ __device__ inline uint32_t pack_half2_hd(float x, float y) {
  return (uint32_t)f32_to_f16_rne(x) | ((uint32_t)f32_to_f16_rne(y) << 16);
}

__host__ __device__ inline float f16_to_f32(uint16_t h) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (float)__builtin_bit_cast(_Float16, h);
#else
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
  union {
    uint32_t u;
    float f;
  } r;
  if (e == 0) {
    const float v = std::ldexp((float)m, -24);
    return sign ? -v : v;
  }
  r.u = (e == 31) ? (sign | 0x7F800000u | (m << 13)) : (sign | ((e + 112u) << 23) | (m << 13));
  return r.f;
#endif
}

__host__ __device__ inline void unpack_half2_hd(uint32_t u, float& x, float& y) {
  x = f16_to_f32((uint16_t)(u & 0xFFFFu));
  y = f16_to_f32((uint16_t)(u >> 16));
}

// RD(m^2), m the midpoint between R >= 0 and the next float: the bound of
// the exact test sqrt_rn(x) > R  <=>  x > RD(m^2) for a float x (see the file
// header). With u = ulp(R), m^2 = R next(R) + u^2 / 4, where R next(R) is an
// integer multiple of u^2 with 47 or 48 significant bits, so u^2 / 4 lies
// below one ulp of it and RD(m^2) = RD(R next(R)): the rounded product, moved
// down one float when its FMA error shows it was rounded up. For R = 0 the
// product and the bound are 0. (Checked against the exact FP64 m^2 for every
// float of [2^-16, 2^5) on the host.)
__host__ __device__ inline float rd_bound(float R) {
  const float n = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, R) + 1u);
  const float p = R * n;
  const float e = __builtin_fmaf(R, n, -p);
  return e < 0.0f ? __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, p) - 1u) : p;
}

// x / b correctly rounded from y = RN(1 / b): q = RN(x y) lies within an ulp
// of x / b, the remainder r = x - q b is exact under an FMA, and RN(q + r y)
// is then RN(x / b) (Markstein's theorem). Holds for every normal x and the
// divisors used here (the integers 1..16 and 3.14), as an exhaustive check
// over all float significands confirms; three operations instead of the
// scaled IEEE division sequence.
__host__ __device__ inline float div_by_recip(float x, float b, float y) {
  const float q = x * y;
  const float r = __builtin_fmaf(-q, b, x);
  return __builtin_fmaf(r, y, q);
}

// RN(1 / w) for w = 0..16 (entry 0 unused); compile-time IEEE divisions.
constexpr float kRcp[17] = {0.0f,        1.0f,        1.0f / 2.0f,  1.0f / 3.0f,  1.0f / 4.0f,
                            1.0f / 5.0f, 1.0f / 6.0f, 1.0f / 7.0f,  1.0f / 8.0f,  1.0f / 9.0f,
                            1.0f / 10.0f, 1.0f / 11.0f, 1.0f / 12.0f, 1.0f / 13.0f, 1.0f / 14.0f,
                            1.0f / 15.0f, 1.0f / 16.0f};
constexpr float kRcp314 = 1.0f / 3.14f;

// Level-0 decode of one atlas texel (pull.comp:37-39, :51-66): depth,
// maxDepth and the radius in ISM pixels. r = float(s & 255) / 10 and
// madd = r * 2 / zFar * 0.6 are passed in.
struct Decoded {
  float depth, max_depth, radius;
};

__host__ __device__ inline Decoded decode_texel(uint32_t s, float r, float madd, float z_far,
                                                int S) {
  Decoded d;
  d.depth = (float)(s >> 8) / 16777216.0f;
  d.max_depth = d.depth + madd;
  const float dist = d.depth * z_far;
#if defined(__HIP_DEVICE_COMPILE__)
  // r / dist from the correctly rounded reciprocal of dist (mlgi_device.hpp;
  // the IEEE quotient's bits). A zero dist gives NaN where the division gives
  // +inf or NaN; both end as the radius 15 of the clamp below.
  const float q = dev::div_rn(r, dist, dev::recip_rn(dist));
#else
  const float q = r / dist;
#endif
  float radius = div_by_recip(q, 3.14f, kRcp314) * (float)S;
  radius = radius * 1.3f;
  d.radius = fminf(radius, 15.0f);
  return d;
}

// pull.comp:84-125: combine four inputs (depth, maxDepth, radius, displacement
// to the output centre, validity) into one output texel.
__host__ __device__ inline Texel pull_combine(const float* ds, const float* mds, const float* rs,
                                              const float* dx, const float* dy, bool* valid) {
  // Branch-free selection: fminf(d, minimum) == d is d <= minimum (no NaNs
  // here). A sum takes each term as fma(v, term, acc) with v = 1 or 0: the
  // term itself, or +-0, which leaves the sum's bits unchanged (a sum starts
  // at +0 and is never -0: an exact cancellation rounds to +0).
  float minimum = __builtin_inff();
  float max_depth = 0.0f;
  bool any = false;
  for (int i = 0; i < 4; ++i) {
    const bool take = valid[i] && !(ds[i] > minimum);
    minimum = take ? ds[i] : minimum;
    max_depth = take ? mds[i] : max_depth;
    any = any || valid[i];
  }
  Texel res{1.0f, 0.0f, 0.0f, 0u};
  if (!any) return res;
  float dacc = 0.0f, racc = 0.0f, xacc = 0.0f, yacc = 0.0f, mvm = 0.0f;
  uint32_t nv = 0;
  for (int i = 0; i < 4; ++i) {
    const bool v = valid[i] && !(ds[i] > max_depth);
    const float vf = v ? 1.0f : 0.0f;
    dacc = __builtin_fmaf(vf, ds[i], dacc);
    xacc = __builtin_fmaf(vf, dx[i], xacc);
    yacc = __builtin_fmaf(vf, dy[i], yacc);
    racc = __builtin_fmaf(vf, rs[i], racc);
    mvm = (v && mvm < mds[i]) ? mds[i] : mvm;
    nv += v ? 1u : 0u;
  }
  if (nv > 0) {
    const float fn = (float)nv;
    const float y = nv == 3 ? kRcp[3] : (nv == 4 ? 0.25f : (nv == 2 ? 0.5f : 1.0f));
    res.depth = div_by_recip(dacc, fn, y);
    res.max_depth = mvm;
    res.radius = div_by_recip(racc, fn, y);
    res.disp = pack_half2_hd(div_by_recip(xacc, fn, y), div_by_recip(yacc, fn, y));
  }
  return res;
}


template <int S>
constexpr int ism_log2() {
  return (S == 8) ? 3 : (S == 16) ? 4 : (S == 32) ? 5 : 6;
}

// unorm16 of a depth in [0, 1] (R16 storage clamps first), as a 32-bit
// word. Every depth here is an atlas depth (s >> 8) / 2^24 or a rounded
// weighted mean of such depths, in [0, 1 + 2^-22] at worst, and anything
// below 1 + 2^-17 rounds to 65535 as 1.0 does, so neither the clamp nor a
// 16-bit mask is needed.
__device__ __forceinline__ uint32_t to_r16(float depth) {
  return (uint32_t)rintf(depth * 65535.0f);
}

// The atlas and the result are read and written through raw buffer
// resources (base and size in scalar registers, a 32-bit byte offset per
// lane: the atlas holds at most 2^28 texels), which spares every access the
// 64-bit address arithmetic of a global pointer. Resource word 3 0x00020000
// selects the 32-bit data format raw buffer accesses use on CDNA; accesses
// past the size read 0 and write nothing.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}

// RN(c + 2 h) for the binary16 h in the low 16 bits of `bits`: doubling the
// widened half is exact, so one FMA (v_fma_mix_f32, which widens the half
// itself) gives the reference's c + h * 2.
__device__ __forceinline__ float plus_twice_half(float c, uint32_t bits) {
  const float h = (float)__builtin_bit_cast(_Float16, (uint16_t)(bits & 0xFFFFu));
  return __builtin_fmaf(h, 2.0f, c);
}

// Pull of output (ox, oy) of level l >= 1 (pull.comp:20-126). Level 1 reads
// the atlas tile at texel `base` of the atlas (row stride 2^stride_log2) and
// decodes it; higher levels read pull level l - 1.
template <int S, int l>
__device__ __forceinline__ Texel pull_one(int ox, int oy, int ism_x, int ism_y, __amdgpu_buffer_rsrc_t atlas,
                                          uint32_t base, int stride_log2, const Texel* src, const float* mtab,
                                          const Decoded& empty, float z_far) {
  constexpr int so = S >> l, si = so * 2;
  constexpr float scale = 1.0f / (float)(1 << l);  // pow(2, -level)
  float ds[4], mds[4], rs[4], dx[4], dy[4];
  bool valid[4];
  // global output position and the first input column / row, as floats
  const float fox = (float)(ism_x * so + ox) + 0.5f, foy = (float)(ism_y * so + oy) + 0.5f;
  const float fix = (float)(2 * (ism_x * so + ox)) + 0.5f, fiy = (float)(2 * (ism_y * so + oy)) + 0.5f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    // offsets {0,0},{0,1},{1,0},{1,1}
    const int offx = i >> 1, offy = i & 1;
    const int ix = ox * 2 + offx, iy = oy * 2 + offy;
    if constexpr (l == 1) {
      const uint32_t s = __builtin_amdgcn_raw_buffer_load_b32(atlas, (base + ((uint32_t)iy << stride_log2) + ix) * 4u, 0, 0);
      const uint32_t b = s & 0xFFu;
      // r = b / 10 (div_by_recip is exact for the divisor 10); the max-depth
      // offset r * 2 / zFar * 0.6 is tabulated per radius byte
      const float r = div_by_recip((float)b, 10.0f, 1.0f / 10.0f);
      const Decoded d = (s == 0xFFFFFFFFu) ? empty : decode_texel(s, r, mtab[b], z_far, S);
      ds[i] = d.depth;
      mds[i] = d.max_depth;
      rs[i] = d.radius;
      // zero displacement: (in + 0.5) / 2 - (out + 0.5) is exactly +-0.25,
      // and sqrt(0.125) rounds to 0.35355339
      dx[i] = offx ? 0.25f : -0.25f;
      dy[i] = offy ? 0.25f : -0.25f;
      valid[i] = ds[i] != 1.0f && 0.35355339f <= rs[i] * scale;
    } else {
      const Texel tx = src[iy * si + ix];
      float disp_x, disp_y;
      unpack_half2_hd(tx.disp, disp_x, disp_y);
      // (in + 0.5 + disp) / 2 - (out + 0.5), in the reference's global texels
      dx[i] = ((fix + (float)offx) + disp_x) * 0.5f - fox;
      dy[i] = ((fiy + (float)offy) + disp_y) * 0.5f - foy;
      ds[i] = tx.depth;
      mds[i] = tx.max_depth;
      rs[i] = tx.radius;
      // sqrt(dx^2 + dy^2) <= radius * scale, decided exactly without the root
      const float x = dx[i] * dx[i] + dy[i] * dy[i];
      valid[i] = ds[i] != 1.0f && !(x > rd_bound(rs[i] * scale));
    }
  }
  return pull_combine(ds, mds, rs, dx, dy, valid);
}

// Push quads wrap around the ISM. The reference's invocations g = 0 .. so/2
// per axis (push.comp:46-67) pair the fine pixels 2g - 1, 2g with the coarse
// neighbours g - 1, g, so the first and the last quad of a row hold one
// in-ISM pixel each, whose other neighbour lies outside the ISM and has
// weight 0 (push.comp:77-86). Here quads g = 0 .. so/2 - 1 take those
// indices mod so (pixels) and mod sc = so/2 (neighbours): quad 0 holds pixel
// so - 1 with neighbour sc - 1 and pixel 0 with neighbour 0, the two border
// invocations in one, and each of its pixels may only use the neighbour on
// its own side. Its pairs across the wrap, which the reference never forms,
// get the offset kAcross instead of +-1.5: the squared distance is then +inf
// and the radius test fails, so they are never valid and never kept, as the
// outside neighbours they stand for. Every quad holds four in-ISM pixels, no
// index is clamped or range-tested, and per 64^2 texels the push levels take
// 16, 4 and 1 wave iterations instead of 18, 5 and 2.
constexpr float kAcross = 1.0e20f;  // absorbs twice any half; its square overflows

// The offset (coarse + 0.5) * 2 - (fine + 0.5) = 2 o - f - 0.5 along one axis
// between the neighbour at o and the pixel at f of a quad (o, f in {0, 1}),
// exact: +-0.5 on the same side, +-1.5 across; kAcross for the pairs across
// the wrap of a quad 0 (`wrapped`).
__device__ __forceinline__ float pair_offset(int o, int f, bool wrapped) {
  const float same = o ? 0.5f : -0.5f, across = wrapped ? kAcross : (o ? 1.5f : -1.5f);
  return o == f ? same : across;
}

// The reference's weightsX = {9, 3, 1, 3} of neighbour i for fine pixel op.
__device__ __forceinline__ float push_weight(int i, int op) {
  const int wt = (i - op + 4) & 3;
  return wt == 0 ? 9.0f : (wt == 2 ? 1.0f : 3.0f);
}

// Offsets {0,0},{0,1},{1,1},{1,0} of a quad's neighbours and of its fine
// pixels, in the reference's invocation order.
constexpr int kOx[4] = {0, 0, 1, 1}, kOy[4] = {0, 1, 1, 0};

// The four coarse neighbours of push quad (gx, gy) of a level l >= 1; their
// usable flags (the neighbour is not empty: depth != 1) are returned apart.
struct Coarse4 {
  float cd[4], cmd[4], cr[4], bound[4];
  uint32_t cdisp[4];
};

template <int S, int l>
__device__ __forceinline__ void load_coarse(int gx, int gy, const Texel* coarse, Coarse4& c, bool (&usable)[4]) {
  constexpr int sc = (S >> l) / 2;
  constexpr float scale = 1.0f / (float)(1 << l);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int cx = (gx - 1 + kOx[i]) & (sc - 1), cy = (gy - 1 + kOy[i]) & (sc - 1);
    const Texel t = coarse[cy * sc + cx];
    c.cd[i] = t.depth;
    c.cmd[i] = t.max_depth;
    c.cr[i] = t.radius;
    c.cdisp[i] = t.disp;
    usable[i] = t.depth != 1.0f;
    c.bound[i] = rd_bound(t.radius * scale);
  }
}

// One pixel of a push invocation (push.comp:44-165) of a level l >= 1: fine
// pixel op (0..3, kOx / kOy) of a quad whose neighbours are c, texel t of
// `same`, which holds its pull value and receives its push value in place.
// wx / wy: the quad is quad 0 of its row / column (the wrap).
// The reference's weights become flags: w_i != 0 is `valid` (usable and
// within the radius), the weight that survives the max-depth test is `keep`.
// The reference sums all four w_i-weighted terms; a dropped term there is
// v * 0 = +-0, and a fill always has a kept term, which is nonzero or +0
// (a displacement c + d with c != 0 that cancels rounds to +0), so summing
// the kept terms over +0 gives the same bits.
template <int l>
__device__ __forceinline__ void push_pixel(const Coarse4& c, const bool (&usable)[4], int op, bool wx, bool wy,
                                           const float* rcp, Texel* same, int t) {
  const int fx = op >> 1, fy = (op ^ (op >> 1)) & 1;  // kOx[op], kOy[op]
  const float* cd = c.cd;
  const float* cmd = c.cmd;
  const float* cr = c.cr;
  float ddx[4], ddy[4];
  bool valid[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    // the pair's offset plus twice the displacement
    ddx[i] = plus_twice_half(pair_offset(kOx[i], fx, wx), c.cdisp[i]);
    ddy[i] = plus_twice_half(pair_offset(kOy[i], fy, wy), c.cdisp[i] >> 16);
    const float x = ddx[i] * ddx[i] + ddy[i] * ddy[i];
    valid[i] = usable[i] && !(x > c.bound[i]);  // sqrt(x) <= radius * 2^-level
  }
  // the reference starts the minimum at 9001; every depth is at most
  // 1 + 2^-22, so the first valid neighbour is always taken
  float minimum = 9001.0f, max_depth = 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {  // fminf(cd, minimum) == cd is cd <= minimum
    const bool take = valid[i] && (i == 0 || !(cd[i] > minimum));
    minimum = take ? cd[i] : minimum;
    max_depth = take ? cmd[i] : max_depth;
  }
  Texel res = same[t];
  bool keep[4], any = false, occluded = false;
#pragma unroll
  for (int i = 0; i < 4; ++i) {  // bitwise: the short-circuit form compiled to more instructions
    keep[i] = valid[i] & !(cd[i] > max_depth);
    any |= keep[i];
    occluded |= keep[i] & (res.depth > cmd[i]);
  }
  if (any && (res.depth == 1.0f || occluded)) {
    // kept terms as fma(k, term, acc), k = 1 or 0 (pull_combine); each
    // weighted term w v is rounded first, as the reference's products are
    float wacc = 0.0f, dacc = 0.0f, macc = 0.0f, racc = 0.0f, xacc = 0.0f, yacc = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float w = push_weight(i, op);
      const float k = keep[i] ? 1.0f : 0.0f;
      wacc = __builtin_fmaf(k, w, wacc);
      dacc = __builtin_fmaf(k, cd[i] * w, dacc);
      macc = __builtin_fmaf(k, cmd[i] * w, macc);
      racc = __builtin_fmaf(k, cr[i] * w, racc);
      xacc = __builtin_fmaf(k, ddx[i] * w, xacc);
      yacc = __builtin_fmaf(k, ddy[i] * w, yacc);
    }
    const float y = rcp[(int)wacc];
    res.depth = div_by_recip(dacc, wacc, y);
    res.max_depth = div_by_recip(macc, wacc, y);
    res.radius = div_by_recip(racc, wacc, y);
    res.disp = pack_half2_hd(div_by_recip(xacc, wacc, y), div_by_recip(yacc, wacc, y));
  }
  // level 1 is read once more, by the depth-only level-0 push, which needs
  // its radius only as the bound of the distance test
  if constexpr (l == 1) res.radius = rd_bound(res.radius);
  same[t] = res;
}

// Push quad (gx, gy) of a level l >= 1: updates, in place, the fine pixels
// 2g - 1 + {0,1}^2 (mod so) of level l from the four coarse neighbours
// g - 1 + {0,1}^2 (mod sc) of level l + 1.
template <int S, int l>
__device__ __forceinline__ void push_quad(int gx, int gy, const Texel* coarse, Texel* same,
                                          const float* rcp) {
  static_assert(l >= 1, "level 0 is push_quad_depth");
  constexpr int so = S >> l;
  const bool wx = gx == 0, wy = gy == 0;
  Coarse4 c;
  bool usable[4];
  load_coarse<S, l>(gx, gy, coarse, c, usable);
#pragma unroll
  for (int op = 0; op < 4; ++op) {
    const int px = (2 * gx - 1 + kOx[op]) & (so - 1), py = (2 * gy - 1 + kOy[op]) & (so - 1);
    push_pixel<l>(c, usable, op, wx, wy, rcp, same, py * so + px);
  }
}

// Pixel op of push quad (gx, gy), for levels whose quads take four lanes
// each: every lane loads the quad's neighbours and pushes one pixel.
template <int S, int l>
__device__ __forceinline__ void push_quad_pixel(int gx, int gy, int op, const Texel* coarse, Texel* same,
                                                const float* rcp) {
  static_assert(l >= 1, "level 0 is push_quad_depth");
  constexpr int so = S >> l;
  const bool wx = gx == 0, wy = gy == 0;
  Coarse4 c;
  bool usable[4];
  load_coarse<S, l>(gx, gy, coarse, c, usable);
  const int px = (2 * gx - 1 + (op >> 1)) & (so - 1), py = (2 * gy - 1 + ((op ^ (op >> 1)) & 1)) & (so - 1);
  push_pixel<l>(c, usable, op, wx, wy, rcp, same, py * so + px);
}

// Level-0 push (push.comp:44-165 at level 0, depth only): quad (gx, gy) writes
// the unorm16 depth of the fine pixels 2g - 1 + {0,1}^2 (mod S) from the
// coarse neighbours g - 1 + {0,1}^2 (mod S/2) of level 1, whose radius slots hold
// rd_bound(radius) (level 0 scales the radius by 2^0). The weights of
// push_quad become per-neighbour flags: w_i != 0 is `valid`, the weight that
// survives the max-depth test is `keep`, and a fill needs an occluding kept
// neighbour, so `wacc > 0 && occluded` is just `occluded`. Sums over the kept
// terms in neighbour order are the sums of push_quad (its zero terms add +0).
// The pixels' original samples come from the ISM's tile at texel `base` of
// the atlas, the results go to the same texels of the result (both of row
// stride 2^stride_log2; atlas sides are powers of two).
template <int S>
__device__ __forceinline__ void push_quad_depth(int gx, int gy, const Texel* coarse, __amdgpu_buffer_rsrc_t atlas,
                                                __amdgpu_buffer_rsrc_t out, uint32_t base, int stride_log2,
                                                const float* rcp) {
  constexpr int so = S, sc = S / 2;
  bool usable[4];
  float cd[4], cd3[4], cd9[4], cmd[4], bnd[4], hx[2][4], hy[2][4];
  const bool wx = gx == 0, wy = gy == 0;  // quads wrap as in push_quad
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int cx = (gx - 1 + kOx[i]) & (sc - 1), cy = (gy - 1 + kOy[i]) & (sc - 1);
    const Texel c = coarse[cy * sc + cx];
    cd[i] = c.depth;
    cd3[i] = c.depth * 3.0f;
    cd9[i] = c.depth * 9.0f;
    cmd[i] = c.max_depth;
    bnd[i] = c.radius;
    usable[i] = c.depth != 1.0f;
    // squared offsets to the pixels at fine offset 0 and 1 along x / y
    const float ex0 = plus_twice_half(pair_offset(kOx[i], 0, wx), c.disp);
    const float ex1 = plus_twice_half(pair_offset(kOx[i], 1, wx), c.disp);
    const float ey0 = plus_twice_half(pair_offset(kOy[i], 0, wy), c.disp >> 16);
    const float ey1 = plus_twice_half(pair_offset(kOy[i], 1, wy), c.disp >> 16);
    hx[0][i] = ex0 * ex0;
    hx[1][i] = ex1 * ex1;
    hy[0][i] = ey0 * ey0;
    hy[1][i] = ey1 * ey1;
  }
#pragma unroll
  for (int op = 0; op < 4; ++op) {
    const int px = (2 * gx - 1 + kOx[op]) & (so - 1), py = (2 * gy - 1 + kOy[op]) & (so - 1);
    bool valid[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) valid[i] = usable[i] && !(hx[kOx[op]][i] + hy[kOy[op]][i] > bnd[i]);
    // the reference starts the minimum at 9001; every depth is at most
    // 1 + 2^-22, so the first valid neighbour is always taken
    float minimum = 9001.0f, max_depth = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool take = valid[i] && (i == 0 || !(cd[i] > minimum));
      minimum = take ? cd[i] : minimum;
      max_depth = take ? cmd[i] : max_depth;
    }
    const uint32_t t = base + ((uint32_t)py << stride_log2) + (uint32_t)px;
    const float orig =
        (float)(__builtin_amdgcn_raw_buffer_load_b32(atlas, t * 4u, 0, 0) >> 8) * (1.0f / 16777216.0f);
    bool occluded = false;
    float wacc = 0.0f, dacc = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int wt = (i - op + 4) & 3;  // weightsX = {9, 3, 1, 3}
      const bool keep = valid[i] && !(cd[i] > max_depth);
      occluded = occluded || (keep && orig > cmd[i]);
      // every term is +0 or positive (depths are >= +0), so the first one
      // is the sum +0 + term itself
      const float w = keep ? (wt == 0 ? 9.0f : (wt == 2 ? 1.0f : 3.0f)) : 0.0f;
      const float d = keep ? (wt == 0 ? cd9[i] : (wt == 2 ? cd[i] : cd3[i])) : 0.0f;
      wacc = i == 0 ? w : wacc + w;
      dacc = i == 0 ? d : dacc + d;
    }
    // rcp[0] = 0 keeps the unused wacc = 0 case finite
    const float depth = occluded ? div_by_recip(dacc, wacc, rcp[(int)wacc]) : orig;
    __builtin_amdgcn_raw_buffer_store_b16((uint16_t)to_r16(depth), out, t * 2u, 0, 0);
  }
}

template <int S>
constexpr int level_offset(int l) {  // texels of pull levels 1..l-1 of one ISM
  int o = 0;
  for (int k = 1; k < l; ++k) o += (S >> k) * (S >> k);
  return o;
}

// ISMs per workgroup: as many S x S ISMs as make 64^2 texels (1 at S = 64,
// 4 at 32, 16 at 16, 64 at 8), so that the small levels of several ISMs share
// wave iterations (a level of one output costs a full wave iteration per
// workgroup). At most the frame's ISM count; both are powers of four, so the
// group divides it.
template <int S>
constexpr int kGroupMax = (64 / S) * (64 / S);

// Pull levels in LDS, level-major: level l of every ISM of the group, then
// level l + 1 (so the levels from l on form one contiguous range). Level l of
// group member g starts at lv_offset<S>(l) + g (S >> l)^2.
template <int S>
constexpr int lv_offset(int l) {
  return kGroupMax<S> * level_offset<S>(l);
}

// The two-pass form of a level needs a list of one 16-bit entry per output
// (pull) or per quad (push) of the group, placed in the levels from `from`
// on, which that level neither reads nor writes. Used where the list fits.
template <int S, int L>
constexpr bool list_fits(int entries, int from) {
  return from <= L && entries * 2 <= (lv_offset<S>(L + 1) - lv_offset<S>(from)) * 16;
}
template <int S, int L>
constexpr bool kCompactPull1 = list_fits<S, L>(kGroupMax<S> * (S / 2) * (S / 2), 2);


// One workgroup's group of ISMs ism0 .. ism0 + groups - 1: member g's atlas
// tile and result at origin(S, g) of the atlas and the result (row stride
// 2^side_log2), its pull levels in lv (lv_offset).
struct Tables {
  __amdgpu_buffer_rsrc_t atlas;  // the atlas (buffer_of)
  __amdgpu_buffer_rsrc_t out;    // the unorm16 result (buffer_of)
  Texel* lv;
  const float* rcp;
  const float* mtab;  // r * 2 / zFar * 0.6 per radius byte (r = b / 10)
  int* counts;  // list length of the two-pass pull level 1
  Decoded empty;
  float z_far;
  int ism0, n1d_log2, groups, side_log2;
  // texel offset of group member g's tile in the atlas
  __device__ uint32_t origin(int S, int g) const {
    const int ism = ism0 + g;
    return ((uint32_t)((ism >> n1d_log2) * S) << side_log2) + (uint32_t)((ism & ((1 << n1d_log2) - 1)) * S);
  }
};

// The group size, a compile-time 1 where a group holds one ISM (S = 64), so
// that the group arithmetic folds away there.
template <int S>
__device__ __forceinline__ int group_count(int groups) {
  return kGroupMax<S> == 1 ? 1 : groups;
}

// Lanes of a wave below this one with their bit of m set.
__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Appends v0 and v1 to an LDS list for the lanes with take0 / take1 set (the
// wave's v0 entries first), with one atomic per wave. Called by every active
// lane of a wave together; in the strided loop below a wave's lane 0 holds
// its lowest index, so it is active whenever any lane is.
__device__ __forceinline__ void list_append2(uint16_t* list, int* count, bool take0, int v0, bool take1, int v1) {
  const unsigned long long m0 = __ballot(take0), m1 = __ballot(take1);
  if ((m0 | m1) == 0ull) return;
  const int n0 = __popcll(m0);
  int first = 0;
  if ((threadIdx.x & 63) == 0) first = atomicAdd(count, n0 + __popcll(m1));
  first = __shfl(first, 0, 64);
  if (take0) list[first + lanes_below(m0)] = (uint16_t)v0;
  if (take1) list[first + n0 + lanes_below(m1)] = (uint16_t)v1;
}

template <int S, int l, int L>
__device__ __forceinline__ void pull_levels(const Tables& tb, const PullConsts& pc) {
  if constexpr (l <= L) {
    constexpr int so = S >> l, n = so * so;
    const int total = group_count<S>(tb.groups) * n;
    auto full = [&](int t) {
      const int g = t / n, o = t - g * n;
      const int ism = tb.ism0 + g;
      const int ism_x = ism & ((1 << tb.n1d_log2) - 1), ism_y = ism >> tb.n1d_log2;
      tb.lv[lv_offset<S>(l) + t] =
          pull_one<S, l>(o % so, o / so, ism_x, ism_y, tb.atlas, tb.origin(S, g), tb.side_log2,
                         tb.lv + lv_offset<S>(l - 1) + g * (4 * n), tb.mtab, tb.empty, tb.z_far);
    };
    if constexpr (l == 1 && kCompactPull1<S, L>) {
      // pass 1: all-empty 2x2 blocks are E_1, the others go to the list.
      // A lane takes two neighbouring blocks of a row (so is even), whose
      // two texel rows it reads with 16-byte loads.
      uint16_t* list = reinterpret_cast<uint16_t*>(tb.lv + lv_offset<S>(2));
      for (int t2 = threadIdx.x; t2 < total / 2; t2 += kThreads) {
        const int t = 2 * t2;  // blocks t and t + 1
        const int g = t / n, o = t - g * n;
        const int ox = o % so, oy = o / so;
        const uint32_t row = (tb.origin(S, g) + ((uint32_t)(2 * oy) << tb.side_log2) + (uint32_t)(2 * ox)) * 4u;
        const auto a = __builtin_amdgcn_raw_buffer_load_b128(tb.atlas, row, 0, 0);
        const auto b = __builtin_amdgcn_raw_buffer_load_b128(tb.atlas, row + (4u << tb.side_log2), 0, 0);
        const bool empty0 = (a[0] & a[1] & b[0] & b[1]) == 0xFFFFFFFFu;
        const bool empty1 = (a[2] & a[3] & b[2] & b[3]) == 0xFFFFFFFFu;
        if (empty0) tb.lv[lv_offset<S>(1) + t] = pc.e1;
        if (empty1) tb.lv[lv_offset<S>(1) + t + 1] = pc.e1;
        list_append2(list, &tb.counts[0], !empty0, t, !empty1, t + 1);
      }
      __syncthreads();
      const int count = tb.counts[0];
      for (int i = threadIdx.x; i < count; i += kThreads) full(list[i]);
    } else {
      for (int t = threadIdx.x; t < total; t += kThreads) full(t);
    }
    __syncthreads();
    pull_levels<S, l + 1, L>(tb, pc);
  }
}

template <int S, int l, int L>
__device__ __forceinline__ void push_levels(const Tables& tb) {
  if constexpr (l >= 0) {
    constexpr int so = S >> l, sc = so / 2, nq1 = so / 2, nq = nq1 * nq1;  // quads wrap (push_quad)
    const int total = group_count<S>(tb.groups) * nq;
    auto full = [&](int t) {
      const int g = t / nq, q = t - g * nq;
      const Texel* coarse = tb.lv + lv_offset<S>(l + 1) + g * (sc * sc);
      if constexpr (l == 0)
        push_quad_depth<S>(q % nq1, q / nq1, coarse, tb.atlas, tb.out, tb.origin(S, g), tb.side_log2, tb.rcp);
      else
        push_quad<S, l>(q % nq1, q / nq1, coarse, tb.lv + lv_offset<S>(l) + g * (so * so), tb.rcp);
    };
    if constexpr (l >= 1 && kGroupMax<S> * nq * 4 <= kThreads) {
      // the level's quads fill at most one workgroup iteration with four
      // lanes each: one pixel per lane. A wave pushes a quad in ~500
      // instructions and a pixel in ~210, and while a small level runs the
      // other waves wait at the barrier, so its critical path counts more
      // than its instruction total (push level 2 at S = 64: 64 quads on one
      // wave, or 256 pixels on four, 1.4 us faster per 1080p frame)
      const int t = threadIdx.x >> 2, op = threadIdx.x & 3;
      if (t < total) {
        const int g = t / nq, q = t - g * nq;
        push_quad_pixel<S, l>(q % nq1, q / nq1, op, tb.lv + lv_offset<S>(l + 1) + g * (sc * sc),
                              tb.lv + lv_offset<S>(l) + g * (so * so), tb.rcp);
      }
    } else {
      for (int t = threadIdx.x; t < total; t += kThreads) full(t);
    }
    __syncthreads();
    push_levels<S, l - 1, L>(tb);
  }
}

// LDS of one workgroup's pull-push: the pull levels of kGroupMax<S> ISMs
// (4096 texels).
template <int S, int L>
struct TileLds {
  static constexpr int kLv = level_offset<S>(L + 1);
  Texel lv[kGroupMax<S> * kLv];
  float mtab[256], rcp[17];
  int counts[1];  // list length of the two-pass pull level 1
};

// Pull-push of ISMs ism0 .. ism0 + groups - 1 of an atlas, by the whole
// workgroup; ends with the unorm16 tiles written (the caller synchronises
// before reusing the LDS).
template <int S, int L>
__device__ __forceinline__ void pullpush_tile(TileLds<S, L>& sm, const uint32_t* __restrict__ atlas,
                                              int atlas_side, float z_far, const PullConsts& pc,
                                              uint16_t* __restrict__ out, int ism0, int groups_) {
  const int tid = threadIdx.x;
  const int groups = group_count<S>(groups_);
  const int n1d_log2 = 31 - __builtin_clz(atlas_side / S);

  // max-depth offsets per radius byte (the reference's expression)
  for (int b = tid; b < 256; b += kThreads) sm.mtab[b] = (float)b / 10.0f * 2.0f / z_far * 0.6f;
  if (tid < 17) sm.rcp[tid] = kRcp[tid];
  if (tid == 0) sm.counts[0] = 0;
  __syncthreads();
  const uint32_t texels = (uint32_t)atlas_side * (uint32_t)atlas_side;
  Tables tb{buffer_of(atlas, texels * 4u), buffer_of(out, texels * 2u), sm.lv, sm.rcp, sm.mtab, sm.counts, {},
            z_far, ism0, n1d_log2, groups, 31 - __builtin_clz(atlas_side)};
  // decode of an empty texel (0xFFFFFFFF), shared by every empty texel
  const float r255 = (float)255 / 10.0f;
  tb.empty = decode_texel(0xFFFFFFFFu, r255, r255 * 2.0f / z_far * 0.6f, z_far, S);

  pull_levels<S, 1, L>(tb, pc);
  if constexpr (L == 1) {
    // no level-1 push stores the bounds the level-0 push reads
    constexpr int n1 = (S / 2) * (S / 2);
    for (int t = tid; t < groups * n1; t += kThreads) {
      Texel& x = sm.lv[t];
      x.radius = rd_bound(x.radius);
    }
    __syncthreads();
  }
  push_levels<S, L - 1, L>(tb);
}

template <int S, int L>
__global__ __launch_bounds__(kThreads) void pullpush_kernel(const uint32_t* __restrict__ atlas,
                                                            int atlas_side, float z_far, PullConsts pc,
                                                            uint16_t* __restrict__ out, int groups) {
  __shared__ TileLds<S, L> sm;
  pullpush_tile<S, L>(sm, atlas, atlas_side, z_far, pc, out, blockIdx.x * groups, groups);
}

// Batched: one workgroup per (frame, ISM group) item of the frames that share
// the ISM side S and the level count L. ISMs are independent, so no
// persistent loop is needed: the hardware keeps as many workgroups resident
// as LDS and registers allow. (A persistent loop over items doubled the
// register count, 105 against 57 at S = 32.)
template <int S, int L>
__global__ __launch_bounds__(kThreads) void batch_pullpush_kernel(const BatchFrame* __restrict__ frames,
                                                                  int n, uint32_t total) {
  __shared__ TileLds<S, L> sm;
  const uint32_t item = blockIdx.x;
  if (item >= total) return;
  const BatchFrame& f = frames[batch_find<&BatchFrame::first_pp>(frames, n, item)];
  PullConsts pc;
  pc.e1 = Texel{__uint_as_float(f.pp_e1[0]), __uint_as_float(f.pp_e1[1]), __uint_as_float(f.pp_e1[2]),
                f.pp_e1[3]};
  const int groups = pullpush_group(S, f.n1d * f.n1d);
  pullpush_tile<S, L>(sm, f.atlas, f.atlas_side, f.z_far, pc, const_cast<uint16_t*>(f.g.ism),
                      (int)(item - f.first_pp) * groups, groups);
}

// Host-side constant: pull level 1 of an all-empty 2x2 block, computed with
// the same functions the kernel uses.
PullConsts make_consts(int S, float z_far) {
  PullConsts pc{};
  const float r = (float)255 / 10.0f;
  const float madd = r * 2.0f / z_far * 0.6f;
  const Decoded e0 = decode_texel(0xFFFFFFFFu, r, madd, z_far, S);
  float ds[4], mds[4], rs[4], dx[4], dy[4];
  bool valid[4];
  for (int i = 0; i < 4; ++i) {
    ds[i] = e0.depth;
    mds[i] = e0.max_depth;
    rs[i] = e0.radius;
    dx[i] = (i >> 1) ? 0.25f : -0.25f;
    dy[i] = (i & 1) ? 0.25f : -0.25f;
    valid[i] = ds[i] != 1.0f && 0.35355339f <= rs[i] * 0.5f;
  }
  pc.e1 = pull_combine(ds, mds, rs, dx, dy, valid);
  return pc;
}

template <int S, int L>
hipError_t launch_sl(hipStream_t s, const uint32_t* atlas, int atlas_side, float z_far,
                     const PullConsts& pc, uint16_t* out) {
  const int n1d = atlas_side / S;
  const int groups = pullpush_group(S, n1d * n1d);
  hipLaunchKernelGGL((pullpush_kernel<S, L>), dim3(n1d * n1d / groups), dim3(kThreads), 0, s, atlas, atlas_side,
                     z_far, pc, out, groups);
  return hipGetLastError();
}

template <int S>
hipError_t launch_s(hipStream_t s, const uint32_t* atlas, int atlas_side, int levels, float z_far,
                    const PullConsts& pc, uint16_t* out) {
  constexpr int kMax = ism_log2<S>();
  switch (levels) {
    case 1: return launch_sl<S, 1>(s, atlas, atlas_side, z_far, pc, out);
    case 2: return launch_sl<S, 2>(s, atlas, atlas_side, z_far, pc, out);
    case 3: return launch_sl<S, 3>(s, atlas, atlas_side, z_far, pc, out);
    case 4:
      if constexpr (kMax >= 4) return launch_sl<S, 4>(s, atlas, atlas_side, z_far, pc, out);
      break;
    case 5:
      if constexpr (kMax >= 5) return launch_sl<S, 5>(s, atlas, atlas_side, z_far, pc, out);
      break;
    case 6:
      if constexpr (kMax >= 6) return launch_sl<S, 6>(s, atlas, atlas_side, z_far, pc, out);
      break;
    default: break;
  }
  return hipErrorInvalidValue;
}

template <int S, int L>
hipError_t launch_batch_sl(hipStream_t s, const BatchFrame* frames, int n, uint32_t total) {
  hipLaunchKernelGGL((batch_pullpush_kernel<S, L>), dim3(total), dim3(kThreads), 0, s,
                     frames, n, total);
  return hipGetLastError();
}

template <int S>
hipError_t launch_batch_s(hipStream_t s, const BatchFrame* frames, int n, uint32_t total, int levels) {
  constexpr int kMax = ism_log2<S>();
  switch (levels) {
    case 1: return launch_batch_sl<S, 1>(s, frames, n, total);
    case 2: return launch_batch_sl<S, 2>(s, frames, n, total);
    case 3: return launch_batch_sl<S, 3>(s, frames, n, total);
    case 4:
      if constexpr (kMax >= 4) return launch_batch_sl<S, 4>(s, frames, n, total);
      break;
    case 5:
      if constexpr (kMax >= 5) return launch_batch_sl<S, 5>(s, frames, n, total);
      break;
    case 6:
      if constexpr (kMax >= 6) return launch_batch_sl<S, 6>(s, frames, n, total);
      break;
    default: break;
  }
  return hipErrorInvalidValue;
}

}  // namespace

void pullpush_empty_texel(int ism_side, float z_far, uint32_t out[4]) {
  const PullConsts pc = make_consts(ism_side, z_far);
  out[0] = f2u(pc.e1.depth);
  out[1] = f2u(pc.e1.max_depth);
  out[2] = f2u(pc.e1.radius);
  out[3] = pc.e1.disp;
}

hipError_t launch_batch_pullpush(hipStream_t s, const BatchFrame* frames, int n, uint32_t total,
                                 int ism_side, int levels) {
  if (total == 0) return hipSuccess;
  switch (ism_side) {
    case 8: return launch_batch_s<8>(s, frames, n, total, levels);
    case 16: return launch_batch_s<16>(s, frames, n, total, levels);
    case 32: return launch_batch_s<32>(s, frames, n, total, levels);
    case 64: return launch_batch_s<64>(s, frames, n, total, levels);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_pullpush(hipStream_t s, const uint32_t* atlas, int atlas_side, int ism_side,
                           int levels, float z_far, uint16_t* out) {
  const PullConsts pc = make_consts(ism_side, z_far);
  switch (ism_side) {
    case 8: return launch_s<8>(s, atlas, atlas_side, levels, z_far, pc, out);
    case 16: return launch_s<16>(s, atlas, atlas_side, levels, z_far, pc, out);
    case 32: return launch_s<32>(s, atlas, atlas_side, levels, z_far, pc, out);
    case 64: return launch_s<64>(s, atlas, atlas_side, levels, z_far, pc, out);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace mlgi
