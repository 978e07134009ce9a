// Interleaved final gathering with per-tile VPL culling.
// Follows gi/final_gathering.comp:63-135 (brute-force class list) and the
// dispatch of GIStage::compute_final_gathering (GIStage.cpp:261-306).
//
// Work decomposition. Pixel (x, y) belongs to interleave class
// k = (x mod I) + (y mod I) * I and sums the VPLs [k*N/I^2, (k+1)*N/I^2).
// A workgroup owns a TW x TH tile of one class's pixels (pixels I apart on
// screen, like the reference's 8x8 workgroups); each wave64 of it owns 64 of
// those pixels, so every lane of a wave walks the same VPL list and each VPL
// record is wave-uniform: it is fetched with scalar loads and used from SGPRs.
//
// Culling (replaces the clustered light lists, light_lists.comp:104-119).
// Each wave reduces the world-space AABB of its pixels; lane j tests VPL j of
// a 64-VPL chunk: if the VPL's tangent plane has the whole box behind it by a
// margin far above FP32 rounding, every pixel of the wave would compute
// angleFactor == 0 for it, so dropping it leaves the sum bit-identical. A
// ballot turns the tests into a 64-bit survivor mask the wave iterates.
//
// Workgroup order is XCD-aware: consecutive workgroups go to different XCDs
// (round robin), so the linear work index is remapped such that the I^2
// classes of one screen tile, which read the same G-buffer cache lines, land
// on one XCD and its L2.
#include "kernels.hpp"

namespace mlgi {
namespace {

constexpr int kXcds = 8;

struct VplRec {
  float px, py, pz, packed;
  float nx, ny, nz, sx;
  float sy, sz, ux, uy;
  float uz, cr, cg, cb;
};

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

struct Pixel {
  int x, y;
  bool valid;
  float depth;
  float fx, fy, fz;  // world position
  float nx, ny, nz;  // decoded face normal
};

// Pixel of lane (lx, ly) of tile (tile_x, tile_y) in class k, and its
// reconstruction: ndc = vec4(uv, depth, 1)*2-1, world = VPinv*ndc / w.
__device__ __forceinline__ Pixel load_pixel(const GatherArgs& a, int cx, int cy, int kx, int ky) {
  Pixel p;
  p.x = cx * a.interleave + kx;
  p.y = cy * a.interleave + ky;
  p.valid = p.x < a.width && p.y < a.height;
  const int sx = p.valid ? p.x : 0, sy = p.valid ? p.y : 0;
  const size_t idx = (size_t)sy * a.width + sx;
  const float d = a.depth[idx];
  p.depth = d;
  const float n0 = a.normal[idx * 3 + 0], n1 = a.normal[idx * 3 + 1], n2 = a.normal[idx * 3 + 2];
  const float u = (float)sx / (float)a.width, v = (float)sy / (float)a.height;
  const float c0 = u * 2.0f - 1.0f, c1 = v * 2.0f - 1.0f, c2 = d * 2.0f - 1.0f;
  const float* m = a.vp_inv.m;
  float r[4];
#pragma unroll
  for (int row = 0; row < 4; ++row)
    r[row] = __builtin_fmaf(m[row], c0,
                            __builtin_fmaf(m[4 + row], c1, __builtin_fmaf(m[8 + row], c2, m[12 + row])));
  p.fx = r[0] / r[3];
  p.fy = r[1] / r[3];
  p.fz = r[2] / r[3];
  p.nx = n0 * 2.0f - 1.0f;
  p.ny = n1 * 2.0f - 1.0f;
  p.nz = n2 * 2.0f - 1.0f;
  return p;
}

// Survivor mask of the 64-VPL chunk starting at VPL `first` (count valid).
__device__ __forceinline__ unsigned long long cull_chunk(const GatherArgs& a, int first, int count,
                                                         float cx, float cy, float cz, float hx,
                                                         float hy, float hz) {
  const int lane = threadIdx.x & 63;
  bool keep = false;
  if (lane < count) {
    const float4* r = reinterpret_cast<const float4*>(a.vpl + (size_t)(first + lane) * 16);
    const float4 r0 = r[0], r1 = r[1];
    const float dx = cx - r0.x, dy = cy - r0.y, dz = cz - r0.z;
    const float ax = fabsf(r1.x), ay = fabsf(r1.y), az = fabsf(r1.z);
    // max over the box of n . (p - p_vpl) = n . (c - p_vpl) + |n| . h
    const float support = r1.x * dx + r1.y * dy + r1.z * dz + (ax * hx + ay * hy + az * hz);
    const float mag = ax * fabsf(dx) + ay * fabsf(dy) + az * fabsf(dz) + (ax * hx + ay * hy + az * hz);
    keep = !(support < -1e-4f * mag - 1e-20f);
  }
  return __ballot(keep);
}

// NEAREST texel of the pair re-decided in FP64 (full rate on CDNA): the
// FP32 local coordinate lies within kTie of a texel edge, closer than its
// rounding error can be trusted. Rebuilds the pixel position, the lookAtRH
// basis and the projection from the FP32 inputs in double.
constexpr float kTie = 1e-4f;

__device__ __forceinline__ void refine_texel(const GatherArgs& a, int x, int y, float depth,
                                             const VplRec& v, float* tx, float* ty) {
  const float* m = a.vp_inv.m;
  const double c0 = (double)x / (double)a.width * 2.0 - 1.0;
  const double c1 = (double)y / (double)a.height * 2.0 - 1.0;
  const double c2 = (double)depth * 2.0 - 1.0;
  double r[4];
  for (int row = 0; row < 4; ++row)
    r[row] = (double)m[row] * c0 + (double)m[4 + row] * c1 + ((double)m[8 + row] * c2 + (double)m[12 + row]);
  const double dx = r[0] / r[3] - (double)v.px;
  const double dy = r[1] / r[3] - (double)v.py;
  const double dz = r[2] / r[3] - (double)v.pz;
  const double dist = sqrt(dx * dx + dy * dy + dz * dz);
  const double nx = v.nx, ny = v.ny, nz = v.nz;
  const double upz = (double)0.01f;  // lookAtRH up = (0, 1, 0.01)
  const double qx = ny * upz - nz, qy = -nx * upz, qz = nx;
  const double ql = sqrt(qx * qx + qy * qy + qz * qz);
  const double sx = qx / ql, sy = qy / ql, sz = qz / ql;
  const double ux = sy * nz - sz * ny, uy = sz * nx - sx * nz, uz = sx * ny - sy * nx;
  const double den = 1.0 + (nx * dx + ny * dy + nz * dz) / dist;
  const double lx = (sx * dx + sy * dy + sz * dz) / dist / den;
  const double ly = (ux * dx + uy * dy + uz * dz) / dist / den;
  const double S = (double)a.ism_side;
  *tx = (float)floor((lx + 1.0) * 0.5 * S);
  *ty = (float)floor((ly + 1.0) * 0.5 * S);
}

template <int TW, int TH, bool kShadow>
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

  // wave-level bounds of the pixels' world positions
  const float inf = __builtin_inff();
  const float mnx = wave_min(p.valid ? p.fx : inf), mxx = wave_max(p.valid ? p.fx : -inf);
  const float mny = wave_min(p.valid ? p.fy : inf), mxy = wave_max(p.valid ? p.fy : -inf);
  const float mnz = wave_min(p.valid ? p.fz : inf), mxz = wave_max(p.valid ? p.fz : -inf);
  if (!(mnx <= mxx)) return;  // no pixel of this wave is on screen (wave-uniform)
  const float bcx = 0.5f * (mnx + mxx), bcy = 0.5f * (mny + mxy), bcz = 0.5f * (mnz + mxz);
  const float bhx = 0.5f * (mxx - mnx) * 1.0001f, bhy = 0.5f * (mxy - mny) * 1.0001f,
              bhz = 0.5f * (mxz - mnz) * 1.0001f;

  const int first = k * a.per_class;
  const float S = (float)a.ism_side;
  const float half_s = 0.5f * S;
  const int amask = a.atlas_side - 1;
  const float inv_zfar = 1.0f / a.z_far;
  const float e0 = 1.0f - 0.02f;  // smoothstep(1.0 - shadowBias, 1.0, .)
  const float inv_width = 1.0f / (1.0f - e0);
  float ar = 0.0f, ag = 0.0f, ab = 0.0f;

  for (int chunk = 0; chunk < a.per_class; chunk += 64) {
    const int count = min(64, a.per_class - chunk);
    unsigned long long mask = cull_chunk(a, first + chunk, count, bcx, bcy, bcz, bhx, bhy, bhz);
    while (mask) {
      const int j = __builtin_ctzll(mask);
      mask &= mask - 1ull;
      const int vi = first + chunk + j;
      const VplRec v = *reinterpret_cast<const VplRec*>(a.vpl + (size_t)vi * 16);
      const float dx = p.fx - v.px, dy = p.fy - v.py, dz = p.fz - v.pz;
      const float d2 = dx * dx + dy * dy + dz * dz;
      const float inv = __builtin_amdgcn_rsqf(d2);
      const float ndot = v.nx * dx + v.ny * dy + v.nz * dz;
      const float cv = ndot * inv;
      const float cp = -(p.nx * dx + p.ny * dy + p.nz * dz) * inv;
      const float angle = fmaxf(cv, 0.0f) * fmaxf(cp, 0.0f);
      if (angle > 0.0f) {
        const float inv2 = inv * inv;
        float g = fminf(angle * (inv2 * inv2), a.clamp_value);
        if (kShadow) {
          // paraboloid_project(diff, dist, n, zFar, vi, n1d, false), evaluated
          // in ISM-local texel coordinates: ((p+1)/2) * S = p*S/2 + S/2
          const float lxv = (v.sx * dx + v.sy * dy + v.sz * dz) * inv;
          const float lyv = (v.ux * dx + v.uy * dy + v.uz * dz) * inv;
          const float rden = __builtin_amdgcn_rcpf(1.0f + cv);  // 1 / (1 - v.z)
          const float cx = __builtin_fmaf(lxv * rden, half_s, half_s);
          const float cy = __builtin_fmaf(lyv * rden, half_s, half_s);
          float tx = floorf(cx), ty = floorf(cy);
          const float rx = cx - tx, ry = cy - ty;
          if (fminf(fminf(rx, 1.0f - rx), fminf(ry, 1.0f - ry)) < kTie)
            refine_texel(a, p.x, p.y, p.depth, v, &tx, &ty);
          const int tile_y = vi / a.n1d;
          const int tile_xi = vi - tile_y * a.n1d;
          const int gx = (tile_xi * a.ism_side + (int)tx) & amask;  // GL_REPEAT wrap
          const int gy = (tile_y * a.ism_side + (int)ty) & amask;
          const float occ = (float)a.ism[(size_t)gy * a.atlas_side + gx] * (1.0f / 65535.0f);
          const float sv = d2 * inv * inv_zfar - occ;
          float t = (1.0f - sv - e0) * inv_width;
          t = fminf(fmaxf(t, 0.0f), 1.0f);
          g *= t * t * (3.0f - 2.0f * t);
        }
        ar = __builtin_fmaf(v.cr, g, ar);
        ag = __builtin_fmaf(v.cg, g, ag);
        ab = __builtin_fmaf(v.cb, g, ab);
      }
    }
  }
  if (p.valid) {
    float* o = a.out + ((size_t)p.y * a.width + p.x) * 3;
    o[0] = ar * a.out_scale;
    o[1] = ag * a.out_scale;
    o[2] = ab * a.out_scale;
  }
}

// Lit-pair counter: pairs (pixel, VPL of its class) with angleFactor > 0.
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
    const VplRec v = *reinterpret_cast<const VplRec*>(a.vpl + (size_t)(first + j) * 16);
    const float dx = p.fx - v.px, dy = p.fy - v.py, dz = p.fz - v.pz;
    const float inv = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
    const float cv = (v.nx * dx + v.ny * dy + v.nz * dz) * inv;
    const float cp = -(p.nx * dx + p.ny * dy + p.nz * dz) * inv;
    lit += (p.valid && fmaxf(cv, 0.0f) * fmaxf(cp, 0.0f) > 0.0f) ? 1u : 0u;
  }
  unsigned long long s = lit;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0)
    __hip_atomic_fetch_add(counter, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int TW, int TH>
hipError_t launch_tiled(hipStream_t s, const GatherArgs& a) {
  const int cw = (a.width + a.interleave - 1) / a.interleave;
  const int ch = (a.height + a.interleave - 1) / a.interleave;
  const int tiles_x = (cw + TW - 1) / TW, tiles_y = (ch + TH - 1) / TH;
  const int num_work = tiles_x * tiles_y * a.interleave * a.interleave;
  const int per_xcd = (num_work + kXcds - 1) / kXcds;
  const dim3 grid(per_xcd * kXcds), block(TW * TH);
  if (a.shadowing)
    hipLaunchKernelGGL((final_gather_kernel<TW, TH, true>), grid, block, 0, s, a, tiles_x, num_work,
                       per_xcd);
  else
    hipLaunchKernelGGL((final_gather_kernel<TW, TH, false>), grid, block, 0, s, a, tiles_x,
                       num_work, per_xcd);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_final_gather(hipStream_t s, const GatherArgs& a, int precision, int tile_w,
                               int tile_h) {
  (void)precision;
  if (tile_w == 8 && tile_h == 8) return launch_tiled<8, 8>(s, a);
  if (tile_w == 16 && tile_h == 16) return launch_tiled<16, 16>(s, a);
  if (tile_w == 32 && tile_h == 8) return launch_tiled<32, 8>(s, a);
  if (tile_w == 16 && tile_h == 4) return launch_tiled<16, 4>(s, a);
  return hipErrorInvalidValue;
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
