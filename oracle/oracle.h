/*
 * Oracle for the many-lights GI hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of the reference shaders, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * Nothing in the product library links, loads or calls this code.
 *
 * Every stage exists in two instantiations of one template:
 *   *_f64  FP64 transcription of the GLSL (the accuracy oracle);
 *   *_f32  FP32 twin with a fixed operation order and no FMA contraction
 *          (the bit-exact oracle for the integer/packed outputs: ISM atlas,
 *          pull-push R16 result, packed VPL normals).
 *
 * Followed files (karyon/many-lights-gi):
 *   data/shaders/common/floatpacking.glsl   packing helpers
 *   data/shaders/common/random.glsl:34-66   Jenkins one-at-a-time hash
 *   data/shaders/ism/ism_utils.glsl:5-40    lookAtRH, paraboloid_project
 *   data/shaders/gi/vpl_processor.comp:41-72
 *   data/shaders/ism/ism.comp:54-139
 *   data/shaders/ism/pull.comp:20-126, push.comp:44-165
 *   data/shaders/gi/final_gathering.comp:63-135 (brute-force light list)
 *   source/mfs-painters/multiframepainter/ImperfectShadowmap.cpp:121-192
 *
 * Parity pinning of each function is described in tests/test_oracle_*.py.
 */
#ifndef MLGI_ORACLE_H
#define MLGI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- packing helpers (floatpacking.glsl) ---- */
uint32_t oracle_pack3sn_f32(float x, float y, float z);
uint32_t oracle_pack3sn_f64(double x, double y, double z);
void oracle_unpack3sn_f32(uint32_t bits, float* out3);
void oracle_unpack3sn_f64(uint32_t bits, double* out3);
uint32_t oracle_pack_unorm4x8(float x, float y, float z, float w);
void oracle_unpack_unorm4x8(uint32_t bits, float* out4);
uint16_t oracle_f32_to_f16(float x);
float oracle_f16_to_f32(uint16_t h);

/* ---- random.glsl ---- */
uint32_t oracle_hash1(uint32_t x);
uint32_t oracle_hash3(uint32_t x, uint32_t y, uint32_t z);
float oracle_random3(float x, float y, float z);

/* ---- ism_utils.glsl ---- */
void oracle_look_at_rh_f64(const double* n3, double* s3, double* u3);
void oracle_look_at_rh_f32(const float* n3, float* s3, float* u3);
void oracle_paraboloid_project_f64(const double* rel3, double dist, const double* n3,
                                   double z_far, int ism_index, int ism_indices_1d,
                                   int preserve_sign, double* out3);
void oracle_paraboloid_project_f32(const float* rel3, float dist, const float* n3,
                                   float z_far, int ism_index, int ism_indices_1d,
                                   int preserve_sign, float* out3);
int oracle_ism_indices_1d(int ism_count);

/* ---- VPL extraction (vpl_processor.comp:41-72) ----
 * rsm_diffuse/normal: w*h*3 floats, rsm_depth: w*h floats, row-major, row 0 = bottom.
 * inv_bvp: column-major 4x4 inverse(bias * P * V).
 * shuffle: n ints (or NULL when shuffleLights is false).
 * Outputs are indexed by the shuffled slot. */
void oracle_vpl_extract_f32(const float* diffuse, const float* normal, const float* depth,
                            int w, int h, const float* inv_bvp, float light_intensity,
                            const int32_t* shuffle, int n,
                            float* pos, float* nrm, uint32_t* packed, float* color);
void oracle_vpl_extract_f64(const float* diffuse, const float* normal, const float* depth,
                            int w, int h, const float* inv_bvp, float light_intensity,
                            const int32_t* shuffle, int n,
                            double* pos, double* nrm, uint32_t* packed, double* color);

/* ---- ISM point splat (ism.comp:54-139) ----
 * points: count*4 floats; w holds packUnorm4x8 bits. bin_offsets: n_bins+1.
 * vpl_pos/vpl_nrm: n_vpl*3 (the unpacked quantised normal).
 * atlas: side*side u32, cleared to 0xFFFFFFFF here. */
void oracle_ism_splat_f32(const float* points, const uint32_t* bin_offsets, int n_bins,
                          const float* vpl_pos, const float* vpl_nrm, int n_vpl,
                          int atlas_side, float z_far, uint32_t* atlas);
void oracle_ism_splat_f64(const float* points, const uint32_t* bin_offsets, int n_bins,
                          const float* vpl_pos, const float* vpl_nrm, int n_vpl,
                          int atlas_side, float z_far, uint32_t* atlas);

/* ---- pull-push (pull.comp, push.comp, ImperfectShadowmap.cpp:121-182) ----
 * out_r16: side*side unorm16. out_depth (f64 only, may be NULL): the push
 * level-0 depth before quantisation. */
void oracle_pullpush_f32(const uint32_t* atlas, int atlas_side, int ism_side, int levels,
                         float z_far, uint16_t* out_r16);
void oracle_pullpush_f64(const uint32_t* atlas, int atlas_side, int ism_side, int levels,
                         float z_far, uint16_t* out_r16, double* out_depth);

/* The FP32 splat and pull-push on `threads` host threads (bins / rows in
 * parallel; the same bits as the single-threaded functions), for the CPU
 * baseline's timing. */
void oracle_ism_splat_mt_f32(const float* points, const uint32_t* bin_offsets, int n_bins,
                             const float* vpl_pos, const float* vpl_nrm, int n_vpl,
                             int atlas_side, float z_far, uint32_t* atlas, int threads);
void oracle_pullpush_mt_f32(const uint32_t* atlas, int atlas_side, int ism_side, int levels,
                            float z_far, uint16_t* out_r16, int threads);

/* ---- final gathering (final_gathering.comp:63-135), brute-force class lists ----
 * depth: w*h window depth; normal: w*h*3 encoded n*0.5+0.5.
 * vp_inv: column-major inverse(P*V) of the camera.
 * vpl_*: n_vpl*3 floats. ism: atlas_side^2 unorm16 (may be NULL if !shadowing).
 * out: w*h*3. lit_pairs: number of (pixel, VPL) pairs with angleFactor > 0.
 * threads: worker threads (>=1). */
void oracle_final_gather_f64(const float* depth, const float* normal, int w, int h,
                             const float* vp_inv, const float* vpl_pos, const float* vpl_nrm,
                             const float* vpl_color, int n_vpl, int vpl_start, int vpl_end,
                             const uint16_t* ism, int atlas_side, float z_far,
                             float gi_intensity, float clamp_value, int interleave,
                             int shadowing, int threads, double* out, uint64_t* lit_pairs);
void oracle_final_gather_f32(const float* depth, const float* normal, int w, int h,
                             const float* vp_inv, const float* vpl_pos, const float* vpl_nrm,
                             const float* vpl_color, int n_vpl, int vpl_start, int vpl_end,
                             const uint16_t* ism, int atlas_side, float z_far,
                             float gi_intensity, float clamp_value, int interleave,
                             int shadowing, int threads, float* out, uint64_t* lit_pairs);

/* Final gathering restricted to a pixel subset (rows [y0, y1)), for bounded
 * CPU baselines. Same semantics as above; out covers the selected rows. */
/* VPL-range forms: scaleISMs / pointsOnlyIntoScaledISMs over the VPLs
 * [start, start + count) (ism.comp:31-40,56-66,124; final_gathering.comp:43-45). */
void oracle_ism_splat_range_f32(const float* points, const uint32_t* bin_offsets, int n_bins,
                                const float* vpl_pos, const float* vpl_nrm, int n_vpl, int atlas_side,
                                float z_far, uint32_t* atlas, int start, int count, int scale_isms,
                                int points_only);
void oracle_final_gather_range_f64(const float* depth, const float* normal, int w, int h,
                                   const float* vp_inv, const float* vpl_pos, const float* vpl_nrm,
                                   const float* vpl_col, int n_vpl, int vpl_start, int vpl_end,
                                   const uint16_t* ism, int atlas_side, float z_far, float gi_intensity,
                                   float clamp_value, int interleave, int shadowing, int threads,
                                   double* out, uint64_t* lit_pairs, int ism_offset, int ism_count);
void oracle_final_gather_rows_f32(const float* depth, const float* normal, int w, int h,
                                  int y0, int y1, const float* vp_inv, const float* vpl_pos,
                                  const float* vpl_nrm, const float* vpl_color, int n_vpl,
                                  int vpl_start, int vpl_end, const uint16_t* ism,
                                  int atlas_side, float z_far, float gi_intensity,
                                  float clamp_value, int interleave, int shadowing,
                                  int threads, float* out, uint64_t* lit_pairs);

/* Geometry-aware separable blur of the irradiance (gi_blur.frag, x pass
 * then y pass); projection16 = the camera's P (column-major). */
void oracle_gi_blur_f32(const float* gi, const float* normal, const float* depth, int w, int h,
                        const float* projection16, float* out);
void oracle_gi_blur_f64(const float* gi, const float* normal, const float* depth, int w, int h,
                        const float* projection16, double* out);

/* Deferred composite (deferredshading.frag:47-81, VSM shadow, Reinhard,
 * gamma 2.2), FP64. specular / occlusion may be NULL (0 / 1). */
void oracle_deferred_composite_f64(const float* depth, const float* normal, const float* diffuse,
                                   const float* specular, const float* gi, const float* occlusion, int w, int h,
                                   const float* vsm, int vw, int vh, const float* vp_inv, const float* blvp,
                                   const float* eye, const float* light_pos, const float* inv_light_dir,
                                   float light_intensity, float exposure, double* out);

#ifdef __cplusplus
}
#endif

#endif
