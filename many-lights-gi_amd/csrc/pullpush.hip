// Pull-push hole filling of the ISM atlas, fused per ISM.
// Follows ism/pull.comp:20-126, ism/push.comp:44-165 and the level loop of
// ImperfectShadowmap::pullpush (ImperfectShadowmap.cpp:121-182).
// Compiled with -ffp-contract=off: the R16 result is bit-exact against the
// FP32 oracle twin.
//
// The reference runs 2L full-atlas dispatches through RGBA32F mip chains.
// Here one 256-lane workgroup owns one S x S ISM: it reads the S^2 u32 atlas
// tile once with 16-byte loads into LDS, builds every pull level in LDS,
// pushes back down in place (push level l overwrites pull level l pixel by
// pixel; each pixel's original sample is read by the lane that writes it),
// and writes the unorm16 result once. 2x2 pull blocks never cross ISMs and
// the push border test gives zero weight to every texel of another ISM, so
// ISMs are independent and nothing outside the tile is read. Zero-weight
// terms contribute +-0 to the push sums, so skipping them is exact.
#include "kernels.hpp"
#include "mlgi_device.hpp"

namespace mlgi {
namespace {

struct __align__(16) Texel {
  float depth, max_depth, radius;
  uint32_t disp;  // packHalf2x16 displacement
};

constexpr int kThreads = 256;

template <int S>
constexpr int level_offset(int l) {  // texels of pull levels 1..l-1
  int o = 0;
  for (int k = 1; k < l; ++k) o += (S >> k) * (S >> k);
  return o;
}

template <int S>
__global__ __launch_bounds__(kThreads) void pullpush_kernel(const uint32_t* __restrict__ atlas,
                                                            int atlas_side, int levels,
                                                            float z_far,
                                                            uint16_t* __restrict__ out) {
  constexpr int kLog = (S == 8) ? 3 : (S == 16) ? 4 : (S == 32) ? 5 : (S == 64) ? 6 : 7;
  __shared__ uint32_t base[S * S];
  __shared__ Texel lv[level_offset<S>(kLog + 1)];

  const int tid = threadIdx.x;
  const int n1d = atlas_side / S;
  const int ism_x = blockIdx.x % n1d;
  const int ism_y = blockIdx.x / n1d;
  const size_t a0 = (size_t)ism_y * S * atlas_side + (size_t)ism_x * S;

  // ---- load the ISM's atlas tile (S*S u32) with 16-byte loads
  for (int i = tid; i < S * S / 4; i += kThreads) {
    const int row = (i * 4) / S, col = (i * 4) % S;
    const uint4 v = *reinterpret_cast<const uint4*>(atlas + a0 + (size_t)row * atlas_side + col);
    *reinterpret_cast<uint4*>(&base[row * S + col]) = v;
  }
  __syncthreads();

  // ---- pull (pull.comp), level l written from level l-1
  for (int l = 1; l <= levels; ++l) {
    const int so = S >> l, si = so * 2;
    const float scale = __builtin_ldexpf(1.0f, -l);  // pow(2, -level)
    Texel* dst = lv + level_offset<S>(l);
    const Texel* src = lv + level_offset<S>(l - 1);
    for (int t = tid; t < so * so; t += kThreads) {
      const int ox = t % so, oy = t / so;
      const int gox = ism_x * so + ox, goy = ism_y * so + oy;  // global output texel
      float ds[4], mds[4], rs[4], dx[4], dy[4];
      bool valid[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        // offsets {0,0},{0,1},{1,0},{1,1}
        const int offx = i >> 1, offy = i & 1;
        const int ix = ox * 2 + offx, iy = oy * 2 + offy;
        float depth, max_depth, radius, disp_x = 0.0f, disp_y = 0.0f;
        if (l == 1) {
          const uint32_t s = base[iy * S + ix];
          depth = (float)(s >> 8) / 16777216.0f;
          radius = (float)(s & 0xFFu) / 10.0f;
          max_depth = depth + radius * 2.0f / z_far * 0.6f;
          const float dist = depth * z_far;
          radius = radius / dist / 3.14f * (float)S;
          radius = radius * 1.3f;
          radius = fminf(radius, 15.0f);
        } else {
          const Texel tx = src[iy * si + ix];
          depth = tx.depth;
          max_depth = tx.max_depth;
          dev::unpack_half2(tx.disp, disp_x, disp_y);
          radius = tx.radius;
        }
        const float gix = (float)(gox * 2 + offx), giy = (float)(goy * 2 + offy);
        const float ndx = (gix + 0.5f + disp_x) / 2.0f - ((float)gox + 0.5f);
        const float ndy = (giy + 0.5f + disp_y) / 2.0f - ((float)goy + 0.5f);
        const float dist = sqrtf(ndx * ndx + ndy * ndy);
        ds[i] = depth;
        mds[i] = max_depth;
        rs[i] = radius;
        dx[i] = ndx;
        dy[i] = ndy;
        valid[i] = depth != 1.0f && dist <= radius * scale;
      }
      float minimum = __builtin_inff();
      float max_depth = 0.0f;
      bool any = false;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (!valid[i]) continue;
        any = true;
        minimum = fminf(ds[i], minimum);
        if (minimum == ds[i]) max_depth = mds[i];
      }
      Texel res{1.0f, 0.0f, 0.0f, 0u};
      if (any) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (ds[i] > max_depth) valid[i] = false;
        float dacc = 0.0f, racc = 0.0f, xacc = 0.0f, yacc = 0.0f, mvm = 0.0f;
        uint32_t nv = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (!valid[i]) continue;
          dacc = dacc + ds[i];
          xacc = xacc + dx[i];
          yacc = yacc + dy[i];
          racc = racc + rs[i];
          mvm = fmaxf(mvm, mds[i]);
          ++nv;
        }
        if (nv > 0) {
          const float fn = (float)nv;
          res.depth = dacc / fn;
          res.max_depth = mvm;
          res.radius = racc / fn;
          res.disp = dev::pack_half2(xacc / fn, yacc / fn);
        }
      }
      dst[oy * so + ox] = res;
    }
    __syncthreads();
  }

  // ---- push (push.comp), level l written from level l+1, in place
  const size_t o0 = a0;  // same tile origin in the R16 result
  for (int l = levels - 1; l >= 0; --l) {
    const int so = S >> l, sc = so / 2;
    const float scale = __builtin_ldexpf(1.0f, -l);
    const Texel* coarse = lv + level_offset<S>(l + 1);
    Texel* same = lv + level_offset<S>(l);
    for (int t = tid; t < so * so; t += kThreads) {
      const int px = t % so, py = t / so;
      // invocation g covers fine pixels 2g-1+{0,1}; offsets {0,0},{0,1},{1,1},{1,0}
      const int gx = (px + 1) >> 1, gy = (py + 1) >> 1;
      const int opx = px + 1 - 2 * gx, opy = py + 1 - 2 * gy;
      const int op = opx ? (opy ? 2 : 3) : (opy ? 1 : 0);
      const int clx = gx - 1, cly = gy - 1;
      float cd[4], cmd[4], cr[4], ddx[4], ddy[4];
      int w[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int offx = (i == 2 || i == 3) ? 1 : 0;
        const int offy = (i == 1 || i == 2) ? 1 : 0;
        const int cx = clx + offx, cy = cly + offy;
        const int wtab = (i - op + 4) & 3;  // weightsX = {9, 3, 1, 3}
        w[i] = (wtab == 0) ? 9 : (wtab == 2 ? 1 : 3);
        cd[i] = cmd[i] = cr[i] = ddx[i] = ddy[i] = 0.0f;
        // ISM border test: coarse neighbours outside this ISM get weight 0
        if (cx < 0 || cy < 0 || cx >= sc || cy >= sc) {
          w[i] = 0;
          continue;
        }
        const Texel c = coarse[cy * sc + cx];
        float cdx, cdy;
        dev::unpack_half2(c.disp, cdx, cdy);
        cd[i] = c.depth;
        cmd[i] = c.max_depth;
        cr[i] = c.radius;
        if (c.depth == 1.0f) w[i] = 0;
        const float ctx = ((float)cx + 0.5f) * 2.0f, cty = ((float)cy + 0.5f) * 2.0f;
        const float ttx = (float)px + 0.5f, tty = (float)py + 0.5f;
        ddx[i] = ctx - ttx + cdx * 2.0f;
        ddy[i] = cty - tty + cdy * 2.0f;
        const float dist = sqrtf(ddx[i] * ddx[i] + ddy[i] * ddy[i]);
        if (dist > c.radius * scale) w[i] = 0;
      }
      float minimum = 9001.0f;
      float max_depth = 0.0f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (w[i] == 0) continue;
        minimum = fminf(cd[i], minimum);
        if (minimum == cd[i]) max_depth = cmd[i];
      }
      int wacc = w[0] + w[1] + w[2] + w[3];
      if (wacc > 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (cd[i] > max_depth) w[i] = 0;
      }
      float dacc = 0.0f, macc = 0.0f, racc = 0.0f, xacc = 0.0f, yacc = 0.0f;
      wacc = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (w[i] == 0) continue;
        const float fw = (float)w[i];
        dacc = dacc + cd[i] * fw;
        macc = macc + cmd[i] * fw;
        racc = racc + cr[i] * fw;
        xacc = xacc + ddx[i] * fw;
        yacc = yacc + ddy[i] * fw;
        wacc += w[i];
      }
      Texel res{0.0f, 0.0f, 0.0f, 0u};
      if (wacc > 0) {
        const float fw = (float)wacc;
        res.depth = dacc / fw;
        res.max_depth = macc / fw;
        res.radius = racc / fw;
        res.disp = dev::pack_half2(xacc / fw, yacc / fw);
      }
      Texel orig;
      if (l == 0) {
        orig = Texel{(float)(base[py * S + px] >> 8) / 16777216.0f, 0.0f, 0.0f, 0u};
      } else {
        orig = same[py * so + px];
      }
      const bool invalid = orig.depth == 1.0f;
      bool occluded = false;
#pragma unroll
      for (int i = 0; i < 4; ++i) occluded = occluded || (w[i] > 0 && orig.depth > cmd[i]);
      if (wacc <= 0 || (!invalid && !occluded)) res = orig;
      if (l == 0) {
        const float r = fminf(fmaxf(res.depth, 0.0f), 1.0f);
        out[o0 + (size_t)py * atlas_side + px] = (uint16_t)rintf(r * 65535.0f);
      } else {
        same[py * so + px] = res;
      }
    }
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_pullpush(hipStream_t s, const uint32_t* atlas, int atlas_side, int ism_side,
                           int levels, float z_far, uint16_t* out) {
  const int n1d = atlas_side / ism_side;
  const dim3 grid(n1d * n1d), block(kThreads);
  switch (ism_side) {
    case 8:
      hipLaunchKernelGGL(pullpush_kernel<8>, grid, block, 0, s, atlas, atlas_side, levels, z_far, out);
      break;
    case 16:
      hipLaunchKernelGGL(pullpush_kernel<16>, grid, block, 0, s, atlas, atlas_side, levels, z_far, out);
      break;
    case 32:
      hipLaunchKernelGGL(pullpush_kernel<32>, grid, block, 0, s, atlas, atlas_side, levels, z_far, out);
      break;
    case 64:
      hipLaunchKernelGGL(pullpush_kernel<64>, grid, block, 0, s, atlas, atlas_side, levels, z_far, out);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mlgi
