// Host-side launch wrappers of the gfx950 kernels. Each wrapper validates
// nothing: the C-ABI layer (capi.cpp) checks shapes before calling, so every
// grid and index below is known to be in bounds of its buffers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlgi {

struct Mat4 {
  float m[16];  // column-major
};

// VPL record: 32 floats (128 bytes) per VPL, see include/mlgi.h.
constexpr int kVplFloats = 32;
constexpr int kVplPos = 0;       // xyz, [3] plane offset n . position
constexpr int kVplNormal = 4;    // xyz, [7] zero-flux flags (bit 0 floats, bit 1 halves)
constexpr int kVplS = 8;         // lookAtRH s row
constexpr int kVplU = 12;        // lookAtRH u row
constexpr int kVplFlux = 16;     // rgb, [19] packed normal bits

// vpl_extract.hip
hipError_t launch_vpl_extract(hipStream_t s, const float* diffuse, const float* normal,
                              const float* depth, int w, int h, const Mat4& inv_bvp,
                              float light_intensity, const int32_t* shuffle, int n,
                              float* records);

// ism_splat.hip
hipError_t launch_atlas_clear(hipStream_t s, uint32_t* atlas, size_t texels);
// Point-major splat. The VPL range (scaleISMs / pointsOnlyIntoScaledISMs):
// with points_only, bin b takes its candidates from the range_count VPLs at
// range_start; VPL v splats into ISM v - ism_offset. The plain splat is
// (0, n_vpl, 0, 0).
hipError_t launch_ism_splat(hipStream_t s, const float* points, const uint32_t* bin_offsets,
                            int n_bins, int n_points, const float* vpl_records, int n_vpl, int n1d,
                            int atlas_side, float z_far, uint32_t* atlas, int block, int range_start,
                            int range_count, int ism_offset, int points_only);
// ISM-major splat: per-point candidate masks, one LDS tile per ISM, then the
// spill list of splats that land outside their ISM's tile. Needs a work area
// of splat_work_bytes(n_points) bytes (no initial contents) and a 16-byte
// aligned atlas; writes every atlas texel, so no clear precedes it.
constexpr uint32_t kSpillCap = 16384;  // spill entries per frame
struct SplatWork {
  uint32_t* state;  // [0] spill count, [1] overflow flag
  uint2* spill;     // (texel, value) x kSpillCap
  uint16_t* mask;   // per point: kept candidates as bits
};
bool ism_major_supported(int ism_side);
size_t splat_work_bytes(uint32_t n_points);
SplatWork splat_work(void* base, uint32_t n_points);
// The static work area of the allocation-free splat (ism_splat.hip), sized
// for kScratchMaxPoints points. splat_scratch_init, called with `device`
// current, returns that device's address of it (resolved once per device).
constexpr uint32_t kScratchMaxPoints = 1u << 23;
hipError_t splat_scratch_init(int device, void** out);
// Adds to *counter the number of set candidate bits over the first n_points
// masks of a work area: the splats (LDS atomic mins) of that ISM-major splat.
hipError_t launch_count_splats(hipStream_t s, const SplatWork& w, uint32_t n_points, unsigned long long* counter);
hipError_t launch_ism_splat_tiled(hipStream_t s, const float* points, const uint32_t* bin_offsets, int n_bins,
                                  int n_points, const float* vpl_records, int n_vpl, int n1d, int ism_side,
                                  int atlas_side, float z_far, uint32_t* atlas, const SplatWork& w);

// pullpush.hip
hipError_t launch_pullpush(hipStream_t s, const uint32_t* atlas, int atlas_side, int ism_side,
                           int levels, float z_far, uint16_t* out);

// final_gather.hip
struct GatherArgs {
  const float* depth;
  const float* normal;
  int width, height;
  Mat4 vp_inv;
  const float* vpl;  // records
  int n_vpl;
  int per_class;     // n_vpl / classes
  int interleave;    // I
  int interleave_log2;
  const uint16_t* ism;
  int atlas_side;
  int ism_side;
  int n1d;
  int n1d_log2;
  float z_far;
  float shadow_k1;   // 1 / (bias * zFar), bias = 0.02 (final_gathering.comp:124)
  float shadow_k2;   // 1 / (bias * 65535): R16 texel value to smoothstep units
  float clamp_value;
  float out_scale;   // giIntensityFactor / vplCount * I^2
  int shadowing;
  int cull;          // 1: per-wave VPL culling; 0: every VPL of the class is evaluated
  int ism_offset;    // VPL v samples ISM v - ism_offset (scaleISMs: the VPL range start)
  float* out;
};
hipError_t launch_final_gather(hipStream_t s, const GatherArgs& a, int precision, int tile_w,
                               int tile_h, int fetch);
hipError_t launch_count_lit_pairs(hipStream_t s, const GatherArgs& a,
                                  unsigned long long* counter);

// blur.hip: geometry-aware separable blur (x then y pass, fused)
hipError_t launch_gi_blur(hipStream_t s, const float* gi, const float* normal, const float* depth, int w, int h,
                          const float* projection16, float* out);

// composite.hip: deferred composite (direct light + VSM shadow + GI ambient,
// Reinhard, gamma 2.2). Matrices column-major; specular / occlusion may be null.
struct CompositeArgs {
  const float* depth;
  const float* normal;
  const float* diffuse;
  const float* specular;
  const float* gi;
  const float* occlusion;
  int w, h;
  const float* vsm;  // moments (E[d], E[d^2]) per texel
  int vw, vh;
  float vp_inv[16], blvp[16];
  float eye[3], light_pos[3], inv_light_dir[3];
  float light_intensity, exposure;
  float* out;
};
hipError_t launch_composite(hipStream_t s, const CompositeArgs& a);

// diag.hip: FP32 FMA rate probe (iters must be a multiple of 4)
uint64_t valu_probe_fmas(int blocks, int iters);
hipError_t launch_valu_probe(hipStream_t s, float* out, int blocks, int iters);
// counts[0..1]: reciprocal and quotient mismatches (zeroed by the caller)
uint64_t division_check_quotients();
hipError_t launch_division_check(hipStream_t s, uint32_t* counts, uint32_t seed);

// raycast.hip: G-buffer (rsm = 0: a = depth, b = normal), RSM (rsm = 1:
// a = diffuse, b = normal, c = depth) or surface (rsm = 2: a = albedo,
// b = VSM moments; either may be null) of a scene uploaded to the device
}  // namespace mlgi
namespace mlgi_scene_impl {
struct SceneView;
struct RayCam;
struct Tri;
}  // namespace mlgi_scene_impl
namespace mlgi {
// points.hip: device point set of a scene (records grouped by bin). Work
// buffers: tmp (4n floats), bin (n), count and cursor (n_bins each).
hipError_t launch_sample_points(hipStream_t s, const mlgi_scene_impl::Tri* tris, const double* cdf, int nt,
                                double total, uint64_t seed, float r_packed, int n_bins, int n, float* tmp, int* bin,
                                uint32_t* count, uint32_t* cursor, float* records, uint32_t* offsets);
hipError_t launch_raycast(hipStream_t st, const mlgi_scene_impl::SceneView& s,
                          const mlgi_scene_impl::RayCam& rc, int w, int h, int rsm, float* a, float* b,
                          float* c);

// ------------------------------------------------------------------ batches
// One frame of a batch as the batched kernels see it. The table
// lives in device memory; every stage's work items of all frames form one
// index space, and first_* is the frame's first item in it (frames of a
// pull-push or gather group are contiguous and counted within their group).
struct BatchFrame {
  // VPL extraction
  const float* rsm_diffuse;
  const float* rsm_normal;
  const float* rsm_depth;
  int rsm_w, rsm_h;
  Mat4 inv_bvp;
  float light_intensity;
  const int32_t* shuffle;  // or null
  uint32_t axis_x, axis_y;
  // splat
  const float* points;
  const uint32_t* bin_offsets;
  uint32_t n_points;
  uint32_t* atlas;
  uint16_t* mask;          // splat work area (SplatWork)
  uint32_t* spill_state;
  uint2* spill;
  int atlas_side, ism_side, n1d, levels;
  float z_far;
  uint32_t pp_e1[4];  // pull level 1 of an all-empty block (bits of depth, maxDepth, radius, disp)
  // final gather (g.vpl = VPL records, g.ism = unorm16 ISM atlas, g.out)
  GatherArgs g;
  int wave_w_log2;  // gather wave strip: 2^wave_w_log2 x 64 / 2^wave_w_log2 class pixels
  int strips_x;     // strips per class row
  // first work item per stage
  uint32_t first_vpl, first_tile, first_bin, first_pp, first_wave;  // first_pp: pull-push ISM groups
};

// Work-item counts of a frame, in each stage's unit.
struct BatchCounts {
  uint32_t vpl, tiles, bins, pp, waves;  // pp: pull-push items (ISM groups, pullpush_group)
};

// ISMs of side S that one pull-push workgroup takes: as many as make 64^2
// texels, at most the frame's n_isms (both powers of four, so it divides).
__host__ __device__ inline int pullpush_group(int ism_side, int n_isms) {
  const int g = (64 / ism_side) * (64 / ism_side);
  return g < n_isms ? g : n_isms;
}
BatchCounts batch_counts(const BatchFrame& f);

hipError_t launch_batch_vpl(hipStream_t s, const BatchFrame* frames, int n, uint32_t total);
// ISM-major splat of the batch: mask (bin items), tiles (ISM items), spill (bin items)
hipError_t launch_batch_mask(hipStream_t s, const BatchFrame* frames, int n, uint32_t total);
// max_side: the largest ISM side of the batch's frames (sizes the tile LDS)
hipError_t launch_batch_tiles(hipStream_t s, const BatchFrame* frames, int n, uint32_t total, int max_side);
hipError_t launch_batch_spill(hipStream_t s, const BatchFrame* frames, int n, uint32_t total);
void pullpush_empty_texel(int ism_side, float z_far, uint32_t out[4]);
// frames[0..n) all have the given ISM side and level count
hipError_t launch_batch_pullpush(hipStream_t s, const BatchFrame* frames, int n, uint32_t total,
                                 int ism_side, int levels);
// frames[0..n) all use the given precision
hipError_t launch_batch_gather(hipStream_t s, const BatchFrame* frames, int n, uint32_t total,
                               int precision);

}  // namespace mlgi
