// ISM point splatting with 32-bit min on packed depth.
// Follows ism/ism.comp:54-139 (usePushPull path) and the atlas clear of
// ImperfectShadowmap.cpp:212. Compiled with -ffp-contract=off: every projected
// value is bit-exact against the FP32 oracle twin, and the atlas texel is the
// minimum over all splats that reach it, which no evaluation order changes.
//
// Point-major form (the reference's decomposition; splat_block 64/128/256):
// the atlas is cleared to 0xFFFFFFFF, then one workgroup per point bin b
// stages the 16 candidate VPLs (b+i) mod N in LDS (1 KiB); each lane owns one
// point per iteration, tests all 16 candidates from broadcast LDS reads, keeps
// the first four that pass both hemisphere tests as a bit mask, and projects
// only those. The projected texel receives a no-return, relaxed, agent-scope
// atomic umin in L2. Each splat is a scattered atomic on a line of an atlas
// far larger than L2, so this form is bound by atomic traffic.
//
// ISM-major form (default), three kernels:
//  1. mask: one workgroup per bin, as above, stores each point's 16-bit set of
//     kept candidates (2 bytes per point) instead of projecting;
//  2. tiles: one workgroup per ISM v builds its S x S tile in LDS. The points
//     splatted into v are those of bins (v - i) mod N whose mask holds bit i
//     (i < 16); each wave scans four of those bins 64 points at a time,
//     queues the hits in LDS (ballot + mbcnt) and projects 64 queued points at
//     once, all onto the one, wave-uniform, VPL v, with LDS atomic umin. The
//     tile is then written with 16-byte stores: that write is the clear.
//     A projection can land outside v's tile (exactly on the tile's far
//     border, or one texel across a border by rounding); such splats go to a
//     spill list of kSpillCap entries;
//  3. spill: after every tile is written, applies the spill list with global
//     atomic umin. If the list overflowed, it instead reruns the point-major
//     splat of the whole frame over the tiles; min is idempotent, so the
//     result is the same.
// Tiles are mapped to XCDs in contiguous ranges, so the bins that
// neighbouring ISMs share are read through one L2.
#include <algorithm>
#include <mutex>

#include "batch_device.hpp"
#include "kernels.hpp"
#include "mlgi_device.hpp"

namespace mlgi {
namespace {

using dev::f3;

constexpr int kMaxTest = 16;    // maxVplTestCount (ism.comp:44)
constexpr int kMaxCollect = 4;  // maxVplCollectCount (ism.comp:49)

__global__ __launch_bounds__(256) void atlas_clear_kernel(uint4* __restrict__ atlas, size_t n4) {
  const uint4 ones = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (size_t)gridDim.x * blockDim.x)
    atlas[i] = ones;
}

struct Candidate {
  float px, py, pz;
  float nx, ny, nz;
  float sx, sy, sz;
  float ux, uy, uz;
  int ism;  // ISM index: the VPL id, minus the range start with scaleISMs
  int tile_x, tile_y;
  int pad;
};

__device__ __forceinline__ Candidate load_candidate(const float* __restrict__ vpl, int idx, int n1d,
                                                    int ism_offset = 0) {
  const float4* r = reinterpret_cast<const float4*>(vpl + (size_t)idx * kVplFloats);
  const float4 r0 = gload4(r + kVplPos / 4), r1 = gload4(r + kVplNormal / 4), r2 = gload4(r + kVplS / 4),
               r3 = gload4(r + kVplU / 4);
  Candidate c;
  c.px = r0.x; c.py = r0.y; c.pz = r0.z;
  c.nx = r1.x; c.ny = r1.y; c.nz = r1.z;
  c.sx = r2.x; c.sy = r2.y; c.sz = r2.z;
  c.ux = r3.x; c.uy = r3.y; c.uz = r3.z;
  c.ism = idx - ism_offset;
  // int(ismIndex / ismIndices1d): exact for |index| < 2^24, truncated toward
  // zero (a negative index, possible with scaleISMs, gives tiles left of or
  // above the atlas)
  c.tile_y = c.ism / n1d;
  c.tile_x = c.ism - c.tile_y * n1d;
  c.pad = 0;
  return c;
}

// Radius byte of a point record's packed w (unorm4x8 .w * 25 is the radius;
// the atlas keeps radius * 10 / sqrt(4)).
__device__ __forceinline__ uint32_t radius_byte_of(float w) {
  const float uw = dev::unorm8(__float_as_uint(w) >> 24);
  const float radius = uw * 25.0f;
  return dev::to_uint(radius * 10.0f / 2.0f);
}

// dot(a, b) < 0 for dot = (a.x b.x + a.y b.y) + a.z b.z in FP32, without
// the last addition: rounding keeps the sign of the exact sum of its two
// operands (a nonzero sum of two floats is a multiple of 2^-149, so it never
// rounds to zero, and an exact 0 rounds to +0), so RN(s + t) < 0 is s < -t.
__device__ __forceinline__ bool dot_below_zero(f3 a, f3 b) {
  const float s = a.x * b.x + a.y * b.y;
  return s < -(a.z * b.z);
}

// Bit i set: candidate i is one of the first kMaxCollect that pass both
// hemisphere tests (ism.comp:80-103).
__device__ __forceinline__ uint32_t keep_mask(const Candidate* cand, float4 rec) {
  const f3 pos = dev::make(rec.x, rec.y, rec.z);
  const uint32_t bits = __float_as_uint(rec.w);
  // unpackUnorm4x8: c / 255
  const float ux = dev::unorm8(bits & 0xFFu);
  const float uy = dev::unorm8((bits >> 8) & 0xFFu);
  const float uz = dev::unorm8((bits >> 16) & 0xFFu);
  const f3 pn3 = dev::make(ux * 2.0f - 1.0f, uy * 2.0f - 1.0f, uz * 2.0f - 1.0f);
  // bit i of `pass`: candidate i passes both tests, shifted in from the
  // last candidate down (one add with carry per candidate)
  uint32_t pass = 0;
#pragma unroll
  for (int i = kMaxTest - 1; i >= 0; --i) {
    const Candidate& c = cand[i];
    const f3 rel = dev::sub(pos, dev::make(c.px, c.py, c.pz));
    const f3 nrel = dev::make(-rel.x, -rel.y, -rel.z);
    const bool cull = dot_below_zero(dev::make(c.nx, c.ny, c.nz), rel) || dot_below_zero(pn3, nrel);
    pass = pass + pass + (cull ? 0u : 1u);
  }
  // the first kMaxCollect of them: clear the lowest set bit kMaxCollect
  // times, and keep what was cleared
  uint32_t rest = pass;
#pragma unroll
  for (int k = 0; k < kMaxCollect; ++k) rest &= rest - 1u;
  return pass ^ rest;
}

// Projection of a point onto candidate c (paraboloid_project with
// preserveSign, ism.comp:105-139): the atlas texel (tx, ty) and the packed
// value. False when the texel lies outside the atlas (imageAtomicMin outside
// the image is a no-op).
__device__ __forceinline__ bool project(const Candidate& c, f3 pos, uint32_t radius_byte, float inv_n1d,
                                        float side, float z_far, uint32_t& tx, uint32_t& ty,
                                        uint32_t& value) {
  const f3 rel = dev::sub(pos, dev::make(c.px, c.py, c.pz));
  const float dist = sqrtf(dev::dot(rel, rel));
  float vx = dev::dot(dev::make(c.sx, c.sy, c.sz), rel);
  float vy = dev::dot(dev::make(c.ux, c.uy, c.uz), rel);
  float vz = dev::dot(dev::make(-c.nx, -c.ny, -c.nz), rel);
  const float sign_v = vz > 0.0f ? 1.0f : (vz < 0.0f ? -1.0f : 0.0f);
  // the six IEEE divisions of the reference as quotients from the
  // correctly rounded reciprocals of their three divisors (dev::div_rn: the
  // same bits; a zero divisor gives NaN where the division gives +-inf or
  // NaN, and both make the texel test below fail; a zero quotient's sign is
  // lost in the additions of 1 that follow each quotient)
  const float rdist = dev::recip_rn(dist);
  vx = dev::div_rn(vx, dist, rdist);
  vy = dev::div_rn(vy, dist, rdist);
  vz = dev::div_rn(vz, dist, rdist);
  vz = 1.0f - vz;
  const float rvz = dev::recip_rn(vz);
  vx = dev::div_rn(vx, vz, rvz);
  vy = dev::div_rn(vy, vz, rvz);
  vz = dev::div_rn(dist, z_far, dev::recip_rn(z_far));
  vz = vz * -sign_v;
  vx = (vx + 1.0f) * 0.5f;
  vy = (vy + 1.0f) * 0.5f;
  vx = (vx + (float)c.tile_x) * inv_n1d;
  vy = (vy + (float)c.tile_y) * inv_n1d;
  vx = vx * side;
  vy = vy * side;
  vz = vz * 16777216.0f;
  value = (dev::to_uint(vz) << 8) | radius_byte;
  // ivec2(v.xy) truncates toward zero
  const float fx = truncf(vx), fy = truncf(vy);
  if (!(fx >= 0.0f && fx < side && fy >= 0.0f && fy < side)) return false;
  tx = (uint32_t)fx;
  ty = (uint32_t)fy;
  return true;
}

// The VPL range of a point-major splat (ism.comp:31-40,56-66,124): with
// points_only, bin b's candidates are ((b + i) mod n_vpl) mod count + start;
// VPL v splats into ISM v - ism_offset (the range start with scaleISMs,
// else 0). The identity range is {0, n_vpl, 0, 0}.
struct SplatRange {
  int start, count, ism_offset, points_only;
};

// Stages the 16 candidates of bin b in LDS (lanes 0..15); the caller
// synchronises.
__device__ __forceinline__ void stage_candidates(Candidate* cand, const float* __restrict__ vpl, int n_vpl,
                                                 int n1d, int b, const SplatRange& r) {
  const int lane = threadIdx.x;
  if (lane < kMaxTest) {
    int idx = (b + lane) % n_vpl;  // index %= totalVplCount
    if (r.points_only) idx = idx % r.count + r.start;
    cand[lane] = load_candidate(vpl, idx, n1d, r.ism_offset);
  }
}

// Point-major splat of bin b of one frame: every lane splats one point per
// pass with global atomics. The first point record is requested before the
// staging barrier and each next one while the current one is processed.
template <int kThreads>
__device__ __forceinline__ void splat_bin(Candidate* cand, const float4* __restrict__ points,
                                          const uint32_t* __restrict__ bin_offsets, uint32_t n_points,
                                          const float* __restrict__ vpl, int n_vpl, int n1d,
                                          int atlas_side, float z_far, uint32_t* __restrict__ atlas,
                                          int b, const SplatRange& range) {
  const int lane = threadIdx.x;
  // clamp to the point count so malformed offsets cannot index past the buffer
  const uint32_t end = min(gload(bin_offsets + b + 1), n_points);
  const uint32_t begin = min(gload(bin_offsets + b), end);
  // loads past the bin read its last point (or point 0; n_points > 0) and
  // are discarded: no select of a record, which would go through scratch
  const uint32_t last = end > 0 ? end - 1 : 0;
  uint32_t p = begin + lane;
  float4 rec = gload4(points + min(p, last));
  stage_candidates(cand, vpl, n_vpl, n1d, b, range);
  __syncthreads();

  const float inv_n1d = 1.0f / (float)n1d;  // exact: n1d is a power of two
  const float side = (float)atlas_side;
  while (p < end) {
    const uint32_t pn = p + kThreads;
    const float4 next = gload4(points + min(pn, last));
    const f3 pos = dev::make(rec.x, rec.y, rec.z);
    const uint32_t rb = radius_byte_of(rec.w);
    uint32_t keep = keep_mask(cand, rec);
    while (keep) {
      const int i = __builtin_ctz(keep);
      keep &= keep - 1u;
      uint32_t tx, ty, value;
      if (!project(cand[i], pos, rb, inv_n1d, side, z_far, tx, ty, value)) continue;
      gatomic_min(atlas + (size_t)ty * (uint32_t)atlas_side + tx, value);
    }
    rec = next;
    p = pn;
  }
}

template <int kThreads>
__global__ __launch_bounds__(kThreads) void ism_splat_kernel(
    const float4* __restrict__ points, const uint32_t* __restrict__ bin_offsets, uint32_t n_points,
    const float* __restrict__ vpl, int n_vpl, int n1d, int atlas_side, float z_far,
    uint32_t* __restrict__ atlas, SplatRange range) {
  __shared__ Candidate cand[kMaxTest];
  splat_bin<kThreads>(cand, points, bin_offsets, n_points, vpl, n_vpl, n1d, atlas_side, z_far, atlas,
                      blockIdx.x, range);
}

// ------------------------------------------------------------ ISM-major form

// Mask pass of bin b. The workgroup of bin 0 also resets the frame's spill
// state (count, overflow) for the tile pass that follows.
__device__ __forceinline__ void mask_bin(Candidate* cand, const float4* __restrict__ points,
                                         const uint32_t* __restrict__ bin_offsets, uint32_t n_points,
                                         const float* __restrict__ vpl, int n_vpl, int n1d,
                                         uint16_t* __restrict__ mask, uint32_t* __restrict__ spill_state,
                                         int b) {
  const int lane = threadIdx.x;
  const uint32_t end = min(gload(bin_offsets + b + 1), n_points);
  const uint32_t begin = min(gload(bin_offsets + b), end);
  const uint32_t last = end > 0 ? end - 1 : 0;  // as in splat_bin
  uint32_t p = begin + lane;
  float4 rec = gload4(points + min(p, last));
  if (b == 0 && lane < 2) gstore(spill_state + lane, 0u);
  stage_candidates(cand, vpl, n_vpl, n1d, b, SplatRange{0, n_vpl, 0, 0});
  __syncthreads();
  while (p < end) {
    const uint32_t pn = p + 256;
    const float4 next = gload4(points + min(pn, last));
    gstore(mask + p, (uint16_t)keep_mask(cand, rec));
    rec = next;
    p = pn;
  }
}

constexpr int kTileMax = 64;                   // largest ISM side of the ISM-major form
constexpr int kBinsPerWave = kMaxTest / 4;     // the 16 candidate bins over 4 waves
// Points per bin per step, and the queue entries per wave: a step's hits
// plus fewer than 64 kept from earlier steps.
template <class QueueT>
constexpr int kStep = sizeof(QueueT) == 2 ? 128 : 192;
template <class QueueT>
constexpr int kQueue = kBinsPerWave * kStep<QueueT> + 64;

// 16-bit queue codes: bin slot t (2 bits) and the point's offset in the bin
// relative to the current step's first point, biased by kCodeBias (14 bits).
// Kept entries are re-coded (offset - kStep) at each step; draining the
// queue every kMaxAge steps keeps every offset >= -kCodeBias.
constexpr uint32_t kCodeBias = 8192;
constexpr uint32_t kMaxAge = kCodeBias / kStep<uint16_t>;
static_assert(kStep<uint16_t> < kCodeBias && kBinsPerWave == 4, "codes hold 2 + 14 bits");

// LDS of the tile pass, sized per launch for the largest ISM side S of its
// items: the S x S tile, then the four waves' point queues. Up to S = 32 the
// queue entries are point indices, 192 points per bin per step (13 KiB; 17
// KiB in all with a 32^2 tile, eight workgroups per CU, the wave limit); at
// S = 64 they are 16-bit codes, 128 points per bin per step (4.5 KiB; 20.5
// KiB in all, seven workgroups per CU). The codes cost a decode per point,
// which measured slower where occupancy does not rise.
extern __shared__ __align__(16) uint32_t tile_lds[];

template <class QueueT>
size_t tile_lds_bytes(int max_side) {
  return (size_t)max_side * max_side * sizeof(uint32_t) + 4 * (size_t)kQueue<QueueT> * sizeof(QueueT);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct TileJob {
  const float4* points;
  const uint32_t* bin_offsets;
  uint32_t n_points;
  const float* vpl;
  int n_vpl, n1d, ism_side, atlas_side;
  float z_far;
  uint32_t* atlas;
  const uint16_t* mask;
  uint32_t* spill_state;
  uint2* spill;
};

// One queued point splatted into ISM c (tile origin ox, oy).
__device__ __forceinline__ void splat_queued(const TileJob& j, uint32_t* tile, const Candidate& c, float4 rec,
                                             uint32_t ox, uint32_t oy, float inv_n1d, float side) {
  uint32_t tx, ty, value;
  if (!project(c, dev::make(rec.x, rec.y, rec.z), radius_byte_of(rec.w), inv_n1d, side, j.z_far, tx, ty, value))
    return;
  const uint32_t lx = tx - ox, ly = ty - oy;  // wraps above S when left of / above the tile
  if (lx < (uint32_t)j.ism_side && ly < (uint32_t)j.ism_side) {
    __hip_atomic_fetch_min(&tile[ly * j.ism_side + lx], value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else {
    const uint32_t k = gatomic_add(j.spill_state, 1u);
    if (k < kSpillCap)
      gstore(reinterpret_cast<uint64_t*>(j.spill + k), (uint64_t)(ty * (uint32_t)j.atlas_side + tx) | ((uint64_t)value << 32));
    else
      gatomic_or(j.spill_state + 1, 1u);
  }
}

// Tile pass of ISM v (all 256 lanes; ends after the tile is written). `tile`
// holds at least S * S words, `queues` 4 * kQueue<QueueT> entries.
template <class QueueT>
__device__ __forceinline__ void splat_tile(const TileJob& j, uint32_t* tile, QueueT* queues, int v) {
  constexpr bool kCoded = sizeof(QueueT) == 2;
  constexpr int kSt = kStep<QueueT>, kLoads = kSt / 64;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int S = j.ism_side;
  const int tile_y = v / j.n1d, tile_x = v - tile_y * j.n1d;
  const int n4 = S * S / 4;
  uint4* t4 = reinterpret_cast<uint4*>(tile);
  const uint4 ones = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
  for (int i = tid; i < n4; i += 256) t4[i] = ones;
  __syncthreads();

  if (v < j.n_vpl && j.n_points > 0) {  // ISMs >= N receive no splats
    const Candidate c = load_candidate(j.vpl, v, j.n1d);
    const float inv_n1d = 1.0f / (float)j.n1d;
    const float side = (float)j.atlas_side;
    const uint32_t ox = (uint32_t)(tile_x * S), oy = (uint32_t)(tile_y * S);
    // Wave w owns bins i = w, w + 4, w + 8, w + 12. Per step it reads the
    // masks of up to kSt points of each of its four bins (4 * kSt / 64
    // loads in flight per lane), queues the hits' point indices in LDS (ballot +
    // mbcnt), then splats the queued points 64 at a time with every lane
    // busy, the next 64 records requested while the current ones project.
    QueueT* q = queues + wave * kQueue<QueueT>;
    uint32_t beg[kBinsPerWave], cnt[kBinsPerWave];
    uint32_t longest = 0;
#pragma unroll
    for (int t = 0; t < kBinsPerWave; ++t) {
      const int i = wave + 4 * t;
      const int b = ((v - i) % j.n_vpl + j.n_vpl) % j.n_vpl;  // (b + i) mod N == v
      const uint32_t end = min(gload(j.bin_offsets + b + 1), j.n_points);
      beg[t] = min(gload(j.bin_offsets + b), end);
      cnt[t] = end - beg[t];
      longest = max(longest, cnt[t]);
    }
    // The queue keeps the hits of a step's last, partial round of 64 for the
    // next step (a partial round costs as much as a full one); the last step
    // drains it, and so does every kMaxAge-th step of a coded queue.
    int carry = 0;  // entries q[0, carry) kept from earlier steps (< 64); wave-uniform
    for (uint32_t s0 = 0; s0 < longest; s0 += kSt) {
      uint32_t m[kBinsPerWave][kLoads];
#pragma unroll
      for (int t = 0; t < kBinsPerWave; ++t)
#pragma unroll
        for (int k = 0; k < kLoads; ++k) {
          const uint32_t o = s0 + (uint32_t)(k * 64 + lane);
          m[t][k] = o < cnt[t] ? (uint32_t)gload(j.mask + beg[t] + o) : 0u;
        }
      int qn = carry;  // wave-uniform
#pragma unroll
      for (int t = 0; t < kBinsPerWave; ++t)
#pragma unroll
        for (int k = 0; k < kLoads; ++k) {
          const bool hit = (m[t][k] >> (wave + 4 * t)) & 1u;
          const uint64_t bal = __ballot(hit);
          if (hit) {
            const int slot = qn + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            const uint32_t o = (uint32_t)(k * 64 + lane);
            q[slot] = (QueueT)(kCoded ? ((uint32_t)t << 14) | (o + kCodeBias) : beg[t] + s0 + o);
          }
          qn += __builtin_popcountll(bal);
        }
      wave_sync();
      // entries splatted this step: the whole rounds of 64, or all of them
      // when the queue drains
      const bool drain = s0 + kSt >= longest || (kCoded && (s0 / kSt) % kMaxAge == kMaxAge - 1);
      const int np = drain ? qn : (qn & ~63);
      if (np > 0) {
        // entry -> point index; lanes past the projected entries read the
        // last of them (a valid point)
        auto point_of = [&](int e) {
          const uint32_t x = q[min(e, np - 1)];
          if (!kCoded) return x;
          const uint32_t t = x >> 14;
          const uint32_t b = t == 0 ? beg[0] : t == 1 ? beg[1] : t == 2 ? beg[2] : beg[3];
          return b + s0 + (x & 0x3FFFu) - kCodeBias;  // offset >= -s0
        };
        float4 rec = gload4(j.points + point_of(lane));
        for (int c0 = 0; c0 < np; c0 += 64) {
          const float4 next = gload4(j.points + point_of(c0 + 64 + lane));
          if (c0 + lane < np) splat_queued(j, tile, c, rec, ox, oy, inv_n1d, side);
          rec = next;
        }
      }
      wave_sync();  // the queue is read before it is rewritten
      {
        // move the remaining entries q[np, qn) (fewer than 64) to the front,
        // codes re-coded relative to the next step
        const int rem = qn - np;
        QueueT v = 0;
        if (lane < rem) v = kCoded ? (QueueT)(q[np + lane] - kSt) : q[np + lane];
        wave_sync();
        if (lane < rem) q[lane] = v;
        wave_sync();
        carry = rem;
      }
    }
  }
  __syncthreads();
  const size_t origin = (size_t)tile_y * S * j.atlas_side + (size_t)tile_x * S;
  for (int i = tid; i < n4; i += 256) {
    const int row = (i * 4) / S, col = (i * 4) % S;
    gstore4(reinterpret_cast<uint4*>(j.atlas + origin + (size_t)row * j.atlas_side + col), t4[i]);
  }
}

// Workgroups of the spill pass.
constexpr unsigned kSpillGrid = 256;

// Spill pass of one frame by the whole grid: its spill entries in a
// grid-stride loop or, after an overflow, its point-major splat with the
// bins in a grid-stride loop. The overflow flag is read once per workgroup,
// so every thread of a workgroup takes the same branch (splat_bin
// synchronises the workgroup).
__device__ __forceinline__ void spill_frame(Candidate* cand, const TileJob& j, int n_bins) {
  const uint32_t overflow = gload(j.spill_state + 1);
  if (overflow) {
    for (int b = blockIdx.x; b < n_bins; b += gridDim.x) {
      splat_bin<256>(cand, j.points, j.bin_offsets, j.n_points, j.vpl, j.n_vpl, j.n1d, j.atlas_side, j.z_far,
                     j.atlas, b, SplatRange{0, j.n_vpl, 0, 0});
      __syncthreads();  // cand is restaged for the next bin
    }
    return;
  }
  const uint32_t n = min(gload(j.spill_state), kSpillCap);
  for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < n; k += gridDim.x * 256u) {
    const uint64_t e = gload(reinterpret_cast<const uint64_t*>(j.spill + k));
    gatomic_min(j.atlas + (uint32_t)e, (uint32_t)(e >> 32));
  }
}

TileJob tile_job(const float* points, const uint32_t* bin_offsets, int n_points, const float* vpl, int n_vpl,
                 int n1d, int ism_side, int atlas_side, float z_far, uint32_t* atlas, const SplatWork& w) {
  return TileJob{reinterpret_cast<const float4*>(points), bin_offsets, (uint32_t)n_points, vpl, n_vpl, n1d,
                 ism_side, atlas_side, z_far, atlas, w.mask, w.state, w.spill};
}

__device__ __forceinline__ TileJob tile_job(const BatchFrame& f) {
  return TileJob{reinterpret_cast<const float4*>(f.points), f.bin_offsets, f.n_points, f.g.vpl, f.g.n_vpl, f.n1d,
                 f.ism_side, f.atlas_side, f.z_far, f.atlas, f.mask, f.spill_state, f.spill};
}

// Single-frame kernels: one workgroup per item (xcd_item), XCD x taking one
// contiguous eighth of the items (a persistent grid walking them measured
// slower: the hardware balances retiring workgroups better).
__global__ __launch_bounds__(256) void ism_mask_kernel(TileJob j, int n_bins) {
  __shared__ Candidate cand[kMaxTest];
  const uint32_t b = xcd_item((uint32_t)n_bins);
  if (b >= (uint32_t)n_bins) return;
  mask_bin(cand, j.points, j.bin_offsets, j.n_points, j.vpl, j.n_vpl, j.n1d, const_cast<uint16_t*>(j.mask),
           j.spill_state, (int)b);
}

template <class QueueT>
__global__ __launch_bounds__(256) void ism_tiles_kernel(TileJob j, int n_tiles) {
  QueueT* queues = reinterpret_cast<QueueT*>(tile_lds + j.ism_side * j.ism_side);
  const uint32_t v = xcd_item((uint32_t)n_tiles);
  if (v >= (uint32_t)n_tiles) return;
  splat_tile(j, tile_lds, queues, (int)v);
}

__global__ __launch_bounds__(256) void ism_spill_kernel(TileJob j, int n_bins) {
  __shared__ Candidate cand[kMaxTest];
  spill_frame(cand, j, n_bins);
}

// Batched kernels: one workgroup per (frame, item) of a stage's item space
// (the spill pass stays persistent); frame f's items start at its first_bin
// / first_tile.
__global__ __launch_bounds__(256) void batch_mask_kernel(const BatchFrame* __restrict__ frames, int n,
                                                         uint32_t total) {
  __shared__ Candidate cand[kMaxTest];
  const uint32_t item = xcd_item(total);
  if (item >= total) return;
  const BatchFrame& f = frames[batch_find<&BatchFrame::first_bin>(frames, n, item)];
  mask_bin(cand, reinterpret_cast<const float4*>(f.points), f.bin_offsets, f.n_points, f.g.vpl, f.g.n_vpl,
           f.n1d, f.mask, f.spill_state, (int)(item - f.first_bin));
}

template <class QueueT>
__global__ __launch_bounds__(256) void batch_tiles_kernel(const BatchFrame* __restrict__ frames, int n,
                                                          uint32_t total, int max_side) {
  QueueT* queues = reinterpret_cast<QueueT*>(tile_lds + max_side * max_side);
  const uint32_t item = xcd_item(total);
  if (item >= total) return;
  const BatchFrame& f = frames[batch_find<&BatchFrame::first_tile>(frames, n, item)];
  splat_tile(tile_job(f), tile_lds, queues, (int)(item - f.first_tile));
}

// Batched spill: the frames one after the other, each by the whole grid.
__global__ __launch_bounds__(256) void batch_spill_kernel(const BatchFrame* __restrict__ frames, int n) {
  __shared__ Candidate cand[kMaxTest];
  for (int f = 0; f < n; ++f) {
    spill_frame(cand, tile_job(frames[f]), frames[f].g.n_vpl);
    __syncthreads();
  }
}


}  // namespace

bool ism_major_supported(int ism_side) { return ism_side >= 4 && ism_side <= kTileMax && ism_side % 4 == 0; }

size_t splat_work_bytes(uint32_t n_points) {
  return 256 + (size_t)kSpillCap * sizeof(uint2) + (((size_t)n_points * 2 + 255) & ~(size_t)255);
}

// Static splat work area of the allocation-free splat entry point: one per
// GPU, loaded with the code object and held until the process ends, sized
// for kScratchMaxPoints points. Host code never passes `scratch` itself (its
// host-side name is a stand-in, not a device address): splat_scratch_init
// resolves each device's address once, and that address is what kernels get.
constexpr size_t kScratchBytes = 256 + (size_t)kSpillCap * sizeof(uint2) + (size_t)kScratchMaxPoints * 2;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex scratch_mu;
void* scratch_ptr[kScratchDevices];  // per device, resolved once
}  // namespace

static_assert(kScratchBytes >= 256 + (size_t)kSpillCap * sizeof(uint2) + (((size_t)kScratchMaxPoints * 2 + 255) &
                                                                           ~(size_t)255),
              "the static area holds the work area of kScratchMaxPoints points");

hipError_t splat_scratch_init(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the current device's copy
    if (e != hipSuccess) return e;
    scratch_ptr[device] = p;
  }
  *out = scratch_ptr[device];
  return hipSuccess;
}

SplatWork splat_work(void* base, uint32_t n_points) {
  (void)n_points;
  char* b = static_cast<char*>(base);
  SplatWork w;
  w.state = reinterpret_cast<uint32_t*>(b);
  w.spill = reinterpret_cast<uint2*>(b + 256);
  w.mask = reinterpret_cast<uint16_t*>(b + 256 + (size_t)kSpillCap * sizeof(uint2));
  return w;
}

hipError_t launch_atlas_clear(hipStream_t s, uint32_t* atlas, size_t texels) {
  const size_t n4 = texels / 4;  // atlas sides are multiples of 4
  size_t blocks = (n4 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(atlas_clear_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                     reinterpret_cast<uint4*>(atlas), n4);
  return hipGetLastError();
}

hipError_t launch_ism_splat(hipStream_t s, const float* points, const uint32_t* bin_offsets,
                            int n_bins, int n_points, const float* vpl_records, int n_vpl, int n1d,
                            int atlas_side, float z_far, uint32_t* atlas, int block, int range_start,
                            int range_count, int ism_offset, int points_only) {
  if (n_bins <= 0) return hipSuccess;
  const float4* pts = reinterpret_cast<const float4*>(points);
  const SplatRange r{range_start, range_count, ism_offset, points_only};
  if (block == 64)
    hipLaunchKernelGGL(ism_splat_kernel<64>, dim3(n_bins), dim3(64), 0, s, pts, bin_offsets,
                       (uint32_t)n_points, vpl_records, n_vpl, n1d, atlas_side, z_far, atlas, r);
  else if (block == 128)
    hipLaunchKernelGGL(ism_splat_kernel<128>, dim3(n_bins), dim3(128), 0, s, pts, bin_offsets,
                       (uint32_t)n_points, vpl_records, n_vpl, n1d, atlas_side, z_far, atlas, r);
  else
    hipLaunchKernelGGL(ism_splat_kernel<256>, dim3(n_bins), dim3(256), 0, s, pts, bin_offsets,
                       (uint32_t)n_points, vpl_records, n_vpl, n1d, atlas_side, z_far, atlas, r);
  return hipGetLastError();
}

hipError_t launch_ism_splat_tiled(hipStream_t s, const float* points, const uint32_t* bin_offsets, int n_bins,
                                  int n_points, const float* vpl_records, int n_vpl, int n1d, int ism_side,
                                  int atlas_side, float z_far, uint32_t* atlas, const SplatWork& w) {
  const TileJob j = tile_job(points, bin_offsets, n_points, vpl_records, n_vpl, n1d, ism_side, atlas_side, z_far,
                             atlas, w);
  const bool any = n_bins > 0 && n_points > 0;
  if (any) {
    hipLaunchKernelGGL(ism_mask_kernel, dim3(xcd_items_grid((uint32_t)n_bins)), dim3(256), 0, s, j, n_bins);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const int n_tiles = n1d * n1d;
  const dim3 grid(xcd_items_grid((uint32_t)n_tiles));
  if (ism_side > 32)
    hipLaunchKernelGGL(ism_tiles_kernel<uint16_t>, grid, dim3(256), tile_lds_bytes<uint16_t>(ism_side), s, j, n_tiles);
  else
    hipLaunchKernelGGL(ism_tiles_kernel<uint32_t>, grid, dim3(256), tile_lds_bytes<uint32_t>(ism_side), s, j, n_tiles);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !any) return e;
  // a few entries in the normal case: one workgroup per CU at most (the
  // overflow fallback then strides over the bins)
  hipLaunchKernelGGL(ism_spill_kernel, dim3(kSpillGrid), dim3(256), 0, s, j, n_bins);
  return hipGetLastError();
}

namespace {
// Sum over points of the number of set bits in each point's candidate mask:
// the (point, ISM) splats of the frame, one LDS atomic min each in the tile
// pass. One wave-level reduction and one global atomic per workgroup.
__global__ __launch_bounds__(256) void count_splats_kernel(const uint16_t* __restrict__ mask, uint32_t n,
                                                           unsigned long long* counter) {
  uint32_t c = 0;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u)
    c += (uint32_t)__builtin_popcount((uint32_t)gload(mask + i));
  unsigned long long v = c;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v) __hip_atomic_fetch_add(counter, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
}  // namespace

hipError_t launch_count_splats(hipStream_t s, const SplatWork& w, uint32_t n_points, unsigned long long* counter) {
  if (n_points == 0) return hipSuccess;
  const uint32_t blocks = std::min<uint32_t>((n_points + 255) / 256, 2048u);
  hipLaunchKernelGGL(count_splats_kernel, dim3(blocks), dim3(256), 0, s, w.mask, n_points, counter);
  return hipGetLastError();
}

hipError_t launch_batch_mask(hipStream_t s, const BatchFrame* frames, int n, uint32_t total) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_mask_kernel, dim3(xcd_items_grid(total)), dim3(256), 0, s, frames, n, total);
  return hipGetLastError();
}

hipError_t launch_batch_tiles(hipStream_t s, const BatchFrame* frames, int n, uint32_t total, int max_side) {
  if (total == 0) return hipSuccess;
  if (max_side < 4 || max_side > kTileMax) return hipErrorInvalidValue;
  const dim3 grid(xcd_items_grid(total));
  if (max_side > 32)
    hipLaunchKernelGGL(batch_tiles_kernel<uint16_t>, grid, dim3(256), tile_lds_bytes<uint16_t>(max_side), s, frames, n,
                       total, max_side);
  else
    hipLaunchKernelGGL(batch_tiles_kernel<uint32_t>, grid, dim3(256), tile_lds_bytes<uint32_t>(max_side), s, frames, n,
                       total, max_side);
  return hipGetLastError();
}

hipError_t launch_batch_spill(hipStream_t s, const BatchFrame* frames, int n, uint32_t total) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_spill_kernel, dim3(kSpillGrid), dim3(256), 0, s, frames, n);
  return hipGetLastError();
}

}  // namespace mlgi
