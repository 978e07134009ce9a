/*
 * mlgi -- many-lights global illumination hot path on MI355X (gfx950).
 *
 * C ABI of libmlgi.so. Plain pointers and sizes only; every device pointer
 * is memory from mlgi_malloc (or any hipMalloc'd memory on the context's
 * device). All compute entry points are asynchronous on the context's HIP
 * stream; mlgi_sync() waits for them. A context is externally synchronised
 * (one host thread at a time). Functions return mlgi_status; on failure
 * mlgi_last_error(ctx) holds a message. No exception crosses this ABI.
 *
 * The stage entry points replace the reference's GL stage classes
 * (karyon/many-lights-gi, source/mfs-painters/multiframepainter/):
 *
 *   mlgi_vpl_extract   VPLProcessor::process            VPLProcessor.h:22,
 *                      VPLProcessor.cpp:73-101, gi/vpl_processor.comp
 *   mlgi_ism_splat     ImperfectShadowmap::render (usePushPull branch:
 *                      atlas clear + ism.comp dispatch)  ImperfectShadowmap.cpp:194-261
 *   mlgi_ism_pullpush  ImperfectShadowmap::pullpush     ImperfectShadowmap.cpp:121-182
 *   mlgi_ism_process   ImperfectShadowmap::process      ImperfectShadowmap.h:30-39
 *   mlgi_final_gather  GIStage::compute_final_gathering GIStage.cpp:261-306,
 *                      gi/final_gathering.comp
 *   mlgi_frame         GIStage::process (VPLP -> ISM -> FG) GIStage.cpp:346-418
 *
 * The clustered light lists (ClusteredShading::process, ClusteredShading.h:30-39)
 * are replaced inside mlgi_final_gather by per-tile culling that is exact
 * with respect to the brute-force interleaved sum.
 */
#ifndef MLGI_H
#define MLGI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MLGI_VERSION 1

typedef enum mlgi_status {
  MLGI_OK = 0,
  MLGI_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad size, inconsistent shapes */
  MLGI_ERR_OUT_OF_MEMORY = 2,    /* device allocation failed */
  MLGI_ERR_HIP = 3,              /* HIP runtime error (message in mlgi_last_error) */
  MLGI_ERR_UNSUPPORTED = 4,      /* valid in the reference, not implemented here */
  MLGI_ERR_NO_DEVICE = 5,        /* no HIP device / device index out of range */
  MLGI_ERR_OUT_OF_BUDGET = 6     /* admission refused by an HBM budget */
} mlgi_status;

typedef enum mlgi_precision {
  /* plain FP32 final gathering; NEAREST texels within 2.5e-4 texel of an
   * edge re-decided exactly (float-float arithmetic, no FP64); rel-L2 <= 1e-4
   * against an FP64 evaluation */
  MLGI_PRECISION_FP32 = 0,
  /* the flux as binary16 (record slots 11 and 15) in f16 x f32 mixed FMAs,
   * NEAREST decided in FP32 alone, per-pixel FP32 sums per 64 VPLs combined
   * with Kahan compensation; geometry, projection and shadow terms FP32;
   * rel-L2 <= 2e-3 */
  MLGI_PRECISION_MIXED = 1
} mlgi_precision;

typedef struct mlgi_context mlgi_context;

/* Per-frame parameters: the GIStage members and uniforms
 * (GIStage.h:52-100, defaults GIStage.cpp:215-232, uniforms GIStage.cpp:280-295,
 * ImperfectShadowmap.cpp:34-35,188-190). */
typedef struct mlgi_params {
  float light_intensity;          /* VPL flux scale (lightIntensity, default 5) */
  int32_t shuffle_lights;         /* permute VPLs with mt19937(1979982) shuffle */
  float gi_intensity_factor;      /* default 3000 */
  float vpl_clamping_value;       /* geometry-term clamp, default 1e-3 */
  int32_t vpl_start_index;        /* vplStartIndex: the VPL range's start */
  int32_t vpl_end_index;          /* vplEndIndex; <=0 means vpl count. The gather
                                     normalises by end - start (all VPLs are gathered) */
  int32_t scale_isms;             /* scaleISMs: the atlas holds one ISM per VPL of the
                                     range, VPL v splatting into and gathered from ISM
                                     v - start (indices outside the grid as GL treats
                                     them: invalid atomics, repeat-wrapped sampling);
                                     single-frame entry points only */
  int32_t points_only_into_scaled_isms; /* pointsOnlyIntoScaledISMs: the points are
                                     binned over the range's count and bin b splats
                                     into VPLs ((b + i) mod n) mod count + start;
                                     single-frame entry points only */
  float tess_level_factor;        /* accepted for interface parity; no effect */
  int32_t use_push_pull;          /* must be 1: the point-splat ISM path */
  int32_t enable_shadowing;       /* ENABLE_SHADOWING */
  int32_t use_interleaving;       /* USE_INTERLEAVING */
  int32_t interleave;             /* interleave side I (power of two, reference: 4) */
  float z_far;                    /* ISM / camera zFar, in [2^-60, 2^60] */
  int32_t atlas_side;             /* ISM atlas side (reference totalIsmPixelSize 2048) */
  int32_t pull_push_levels;       /* pull levels L (reference: 6) */
  int32_t precision;              /* mlgi_precision */
  int32_t tile_w, tile_h;         /* final-gather tile in interleave-class pixels; 0 = auto */
  int32_t gather_fetch;           /* final-gather VPL fetch variant 0..1; -1 = auto */
  int32_t splat_block;            /* 0 = ISM-major splat (LDS tiles, no clear);
                                     64/128/256 = point-major splat (global atomics
                                     into a cleared atlas) with that workgroup size */
  int32_t disable_culling;        /* 0 = per-wave VPL culling (default); 1 = every VPL
                                     of the pixel's class is evaluated, as the
                                     reference's full light lists: the same sums,
                                     bit for bit (a verification switch) */
} mlgi_params;

/* RSM inputs (RasterizationStage outputs bound in VPLProcessor.cpp:84-86).
 * Device pointers, row-major, row 0 = bottom (GL texel order). */
typedef struct mlgi_rsm {
  const float* diffuse;     /* width*height*3, RGB */
  const float* face_normal; /* width*height*3, encoded n*0.5+0.5 */
  const float* depth;       /* width*height, window-space depth in [0,1] */
  int32_t width, height;
  float inv_biased_view_projection[16]; /* column-major inverse(bias*P*V) */
} mlgi_rsm;

/* VPL records, MLGI_VPL_RECORD_FLOATS floats (128 bytes) per VPL:
 *  [0..2]   world position          [3]  plane offset n . position
 *  [4..6]   unpacked normal         [7]  zero-flux flags (u32): bit 0 when the
 *                                        flux floats are all +-0, bit 1 when
 *                                        its halves (11, 15) are all zero
 *  [8..10]  lookAtRH s row          [11] flux r | g << 16 as binary16 bits
 *  [12..14] lookAtRH u row          [15] flux b as binary16 bits (high half 0)
 *  [16..18] flux (diffuse * lightIntensity)
 *                                   [19] packed normal (pack3SNToFloat bits, u32)
 * All other slots are 0. Slots 11 and 15 hold [16..18] rounded to nearest
 * even; the mixed-form gather reads the flux from them, and the gather's
 * culling reads slot 7 (records built by hand must fill 7, 11 and 15).
 * (VPLProcessor vplBuffer + packedVplBuffer, VPLProcessor.cpp:29-41). */
#define MLGI_VPL_RECORD_FLOATS 32
typedef struct mlgi_vpls {
  float* records; /* device, count * MLGI_VPL_RECORD_FLOATS floats */
  int32_t count;
} mlgi_vpls;

/* Splat points grouped into bins (one bin per VPL, ism.geom:56-68). Each
 * point is 4 floats: world xyz and, in the 4th slot, the u32 bits of
 * packUnorm4x8(normal*0.5+0.5, radius/25). Bin b holds points
 * [bin_offsets[b], bin_offsets[b+1]). */
typedef struct mlgi_points {
  const float* records;         /* device, count*4 */
  const uint32_t* bin_offsets;  /* device, n_bins+1 */
  int32_t count;
  int32_t n_bins;
} mlgi_points;

/* ISM atlas ("ISM softrender", R32UI): packed (depth24 << 8) | radius8. */
typedef struct mlgi_atlas {
  uint32_t* texels;  /* device, side*side */
  int32_t side;      /* totalIsmPixelSize, power of two */
  int32_t ism_side;  /* ismPixelSize = side / mlgi_ism_indices_1d(ism count) */
} mlgi_atlas;

/* Pull-push result ("Pushpull result", R16 unorm). */
typedef struct mlgi_ism_depth {
  uint16_t* texels;  /* device, side*side */
  int32_t side;
  int32_t ism_side;
} mlgi_ism_depth;

/* Camera G-buffer (faceNormalBuffer, depthBuffer; MultiFramePainter.cpp:118-119). */
typedef struct mlgi_gbuffer {
  const float* depth;       /* width*height window depth */
  const float* face_normal; /* width*height*3, encoded n*0.5+0.5 */
  int32_t width, height;
  float view_projection_inverse[16]; /* column-major viewInverted()*projectionInverted() */
} mlgi_gbuffer;

/* Irradiance output (giBuffer), FP32 RGB. */
typedef struct mlgi_image {
  float* rgb; /* device, width*height*3 */
  int32_t width, height;
} mlgi_image;

/* Per-stage device times of the last profiled frame(s), milliseconds
 * (stage names of the reference's PerfCounter: VPLP, ISM CS, pull-push, FG). */
typedef struct mlgi_stage_times {
  float vpl_ms;
  float ism_clear_ms;
  float ism_splat_ms;
  float pullpush_ms;
  float final_gather_ms;
  float frame_ms;
  int32_t frames; /* frames accumulated since the last reset */
  float blur_ms;  /* mlgi_gi_blur */
} mlgi_stage_times;

/* ---------------------------------------------------------------- context */
int32_t mlgi_version(void);
int mlgi_device_count(int32_t* count);
int mlgi_context_create(int32_t device, mlgi_context** out);
void mlgi_context_destroy(mlgi_context* ctx);
const char* mlgi_last_error(const mlgi_context* ctx);
int mlgi_sync(mlgi_context* ctx);
/* hipStream_t of the context, as an opaque pointer. */
void* mlgi_stream(mlgi_context* ctx);
/* Orders ctx's stream after everything enqueued so far on `after`'s stream
 * (an event recorded on one and waited on by the other; no host thread
 * blocks). Both contexts on one device, neither in use by another thread
 * meanwhile. No counterpart in the reference (it renders on one GL context):
 * lets two contexts pipeline work, e.g. the next frame's RSM render beside
 * the current frame's GI. */
int mlgi_context_wait(mlgi_context* ctx, mlgi_context* after);

/* Defaults of GIStage::initialize (GIStage.cpp:215-232) plus the atlas and
 * pull-push constants of ImperfectShadowmap (2048 side, 6 levels). */
void mlgi_params_default(mlgi_params* p);

/* ----------------------------------------------------------------- memory */
int mlgi_malloc(mlgi_context* ctx, size_t bytes, void** dptr);
int mlgi_free(mlgi_context* ctx, void* dptr);
int mlgi_memcpy_htod(mlgi_context* ctx, void* dst, const void* src, size_t bytes);
int mlgi_memcpy_dtoh(mlgi_context* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy, asynchronous on the context stream. */
int mlgi_memcpy_dtod(mlgi_context* ctx, void* dst, const void* src, size_t bytes);

/* Measured FP32 vector FMA rate of the context's device (lane FMAs per
 * second; one FMA = 2 flops), from a probe kernel of independent FMA chains
 * timed with HIP events (waits for the stream). A measurement aid for the
 * VALU rooflines; no reference counterpart. */
int mlgi_probe_valu(mlgi_context* ctx, double* fma_per_s);

/* Exactness check of the division form the splat's projection uses in place
 * of IEEE division (csrc/mlgi_device.hpp: recip_rn, div_rn): the device
 * reciprocal refined by one Newton step against the IEEE 1 / b for every
 * float significand at eleven exponents (2^-100 .. 2^100), and the quotient
 * from it against the IEEE a / b for four numerators per divisor (random
 * significand and sign, 2^-40 .. 2^40 times the divisor's scale and within
 * 2^-100 .. 2^100; `seed` picks them). Writes both mismatch counts (0 and 0
 * on a device where the splat's atlas is exact) and the number of quotients
 * tested; waits for the stream. A diagnostic; no reference counterpart. */
int mlgi_check_division(mlgi_context* ctx, uint32_t seed, uint64_t* reciprocal_mismatches,
                        uint64_t* quotient_mismatches, uint64_t* quotients_tested);

/* ----------------------------------------------------------------- stages */
/* VPL extraction: count = n VPLs regularly sampled from the RSM. */
int mlgi_vpl_extract(mlgi_context* ctx, const mlgi_rsm* rsm, const mlgi_params* p,
                     mlgi_vpls* out);
/* Clear the atlas to 0xFFFFFFFF and splat every point into <= 4 of the 16
 * candidate ISMs of its bin with 32-bit min. Bit-exact. With splat_block 0
 * (and ISM sides 4..64, a 16-byte aligned atlas) the splat is ISM-major: the
 * context keeps a work area of 2 bytes per point plus 128 KiB, grown on
 * demand (growing it drains the context stream). */
int mlgi_ism_splat(mlgi_context* ctx, const mlgi_points* pts, const mlgi_vpls* vpls,
                   const mlgi_params* p, mlgi_atlas* atlas);
/* mlgi_ism_splat that allocates nothing: no hipMalloc, hipFree, stream
 * synchronisation or profiling event. No counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * kernels and bit-exact atlas as mlgi_ism_splat; the splat's work area
 * (candidate masks, spill list and counters) is a static array in device
 * memory, one per GPU, held until the process ends (16.1 MiB), sized for at
 * most MLGI_NOALLOC_MAX_POINTS points: a larger point set is rejected with
 * MLGI_ERR_INVALID_ARGUMENT. Every call on a GPU shares that GPU's array;
 * the kernels reset what must start at zero (spill count and overflow flag)
 * and write every mask before reading it, so consecutive calls need no
 * clearing, but calls that overlap on one GPU (two contexts or streams of
 * the same device) race and must be ordered by the caller. */
#define MLGI_NOALLOC_MAX_POINTS 8388608
int mlgi_ism_splat_noalloc(mlgi_context* ctx, const mlgi_points* pts, const mlgi_vpls* vpls,
                           const mlgi_params* p, mlgi_atlas* atlas);
/* Diagnostics of the context's last ISM-major splat (waits for the stream):
 * splats that landed outside their ISM's tile, and whether they exceeded the
 * spill list (then the point-major splat ran over the tiles). No reference
 * counterpart. */
int mlgi_splat_spill_stats(mlgi_context* ctx, uint32_t* spills, int32_t* overflowed);
/* Number of (point, ISM) splats of the context's last ISM-major splat run
 * through mlgi_ism_splat (the LDS atomic mins of its tile pass, at most 4
 * per point); waits for the stream. A measurement aid for the splat's
 * atomic rate; no reference counterpart. */
int mlgi_splat_count(mlgi_context* ctx, uint64_t* splats);
/* Pull-push hole filling, one workgroup per ISM, R16 result. Bit-exact.
 * atlas->ism_side (ismPixelSize) must be one of 8, 16, 32, 64; the atlas
 * must be 16-byte and the output 8-byte aligned (as mlgi_malloc returns),
 * and the output must not overlap the atlas (MLGI_ERR_INVALID_ARGUMENT). */
int mlgi_ism_pullpush(mlgi_context* ctx, const mlgi_atlas* atlas, const mlgi_params* p,
                      mlgi_ism_depth* out);
/* Splat + pull-push (ImperfectShadowmap::process). */
int mlgi_ism_process(mlgi_context* ctx, const mlgi_points* pts, const mlgi_vpls* vpls,
                     const mlgi_params* p, mlgi_atlas* atlas, mlgi_ism_depth* out);
/* Interleaved final gathering with per-tile VPL culling. */
int mlgi_final_gather(mlgi_context* ctx, const mlgi_gbuffer* gb, const mlgi_vpls* vpls,
                      const mlgi_ism_depth* ism, const mlgi_params* p, mlgi_image* out);
/* One GI frame: VPL extraction, ISM splat, pull-push, final gathering. The
 * work buffers (vpls, atlas, ism) are caller-owned. */
int mlgi_frame(mlgi_context* ctx, const mlgi_rsm* rsm, const mlgi_points* pts,
               const mlgi_gbuffer* gb, const mlgi_params* p, mlgi_vpls* vpls,
               mlgi_atlas* atlas, mlgi_ism_depth* ism, mlgi_image* out);

/* Geometry-aware separable blur of the irradiance (GIStage::blur,
 * GIStage.cpp:308-344; gi/gi_blur.frag): x pass then y pass with weights
 * from the G-buffer's face normals and linear depth. projection16 is the
 * camera's P (column-major; only P[2][2] and P[3][2] are used). gi and out
 * must be distinct buffers of the G-buffer's size. Asynchronous. */
int mlgi_gi_blur(mlgi_context* ctx, const mlgi_gbuffer* gb, const float* projection16,
                 const mlgi_image* gi, mlgi_image* out);

/* Deferred composite (DeferredShadingStage::process,
 * DeferredShadingStage.cpp:76-131; deferredshading.frag:47-81): direct light
 * with a variance shadow map, GI ambient (gi * diffuse * occlusion),
 * specular, Reinhard tone mapping (exposure) and gamma 2.2. Device buffers of
 * width*height texels (row 0 = bottom); the VSM holds the moments
 * (E[d], E[d^2]) of the light distance per texel, sampled bilinearly with a
 * (0, 0) border. specular and occlusion may be NULL (read as 0 and 1).
 * Matrices are column-major. Asynchronous. */
typedef struct mlgi_composite_inputs {
  const float* depth;      /* window depth */
  const float* normal;     /* encoded n*0.5+0.5, 3 floats */
  const float* diffuse;    /* sRGB albedo, 3 floats */
  const float* specular;   /* sRGB, 3 floats, or NULL */
  const float* gi;         /* irradiance, 3 floats */
  const float* occlusion;  /* ambient occlusion factor, or NULL */
  int32_t width, height;
  const float* vsm;        /* vsm_width*vsm_height*2 moments */
  int32_t vsm_width, vsm_height;
} mlgi_composite_inputs;
typedef struct mlgi_composite_params {
  float view_projection_inverse[16];   /* inverse(P*V) of the camera */
  float biased_light_view_projection[16];
  float camera_eye[3];
  float light_position[3];
  float inverse_light_direction[3];    /* -normalize(lightDirection) */
  float light_intensity;
  float exposure;                      /* default 1 */
} mlgi_composite_params;
int mlgi_deferred_composite(mlgi_context* ctx, const mlgi_composite_inputs* in,
                            const mlgi_composite_params* p, mlgi_image* out);

/* ------------------------------------------------------------ HIP graphs */
typedef struct mlgi_graph mlgi_graph;
/* Capture one mlgi_frame with these descriptors into an instantiated HIP
 * graph (kernel nodes only). Replays read the buffers' current contents, so
 * new RSM / G-buffer data can be uploaded between replays. The graph owns
 * its splat work area (2 bytes per point + 128 KiB, allocated here, freed by
 * mlgi_graph_destroy), so eager work on the context between replays, of any
 * size, never touches memory a replay uses. */
int mlgi_graph_capture_frame(mlgi_context* ctx, const mlgi_rsm* rsm, const mlgi_points* pts,
                             const mlgi_gbuffer* gb, const mlgi_params* p, mlgi_vpls* vpls,
                             mlgi_atlas* atlas, mlgi_ism_depth* ism, mlgi_image* out,
                             mlgi_graph** graph);
/* Enqueue one replay on the context stream (asynchronous). */
int mlgi_graph_launch(mlgi_context* ctx, mlgi_graph* graph);
void mlgi_graph_destroy(mlgi_graph* graph);

/* ---------------------------------------------------------------- batches
 * No counterpart in the reference (it renders one view per frame): many
 * independent GI frames (views or light poses, each with its own shapes and
 * parameters) run as one set of launches, one per stage over the work items
 * of all frames (pull-push and gather once per ISM shape / precision group). Each frame's results are
 * bit-identical to running it alone through the batch path. */
typedef struct mlgi_frame_desc {
  mlgi_rsm rsm;
  mlgi_points points;
  mlgi_gbuffer gbuffer;
  mlgi_params params;   /* per frame; tile_w selects the gather strip width (8, 16, 32) */
  mlgi_vpls vpls;       /* caller-owned work and output buffers */
  mlgi_atlas atlas;
  mlgi_ism_depth ism;   /* must not overlap the atlas, as for mlgi_ism_pullpush */
  mlgi_image image;
} mlgi_frame_desc;
typedef struct mlgi_batch mlgi_batch;
/* Validate every frame and upload the batch's frame table (allocates). */
int mlgi_batch_create(mlgi_context* ctx, const mlgi_frame_desc* frames, int32_t count,
                      mlgi_batch** out);
/* Enqueue all frames (asynchronous). */
int mlgi_batch_run(mlgi_context* ctx, mlgi_batch* batch);
/* Replace frame i's inverse(bias*P*V) in the device frame table
 * (stream-ordered), e.g. after rendering a new light pose into the frame's
 * RSM buffers; runs and graph replays after this call use the new matrix. */
int mlgi_batch_update_light(mlgi_context* ctx, mlgi_batch* batch, int32_t frame,
                            const float* inv_bvp16);
/* Capture one mlgi_batch_run into an instantiated HIP graph. */
int mlgi_batch_capture(mlgi_context* ctx, mlgi_batch* batch, mlgi_graph** graph);
void mlgi_batch_destroy(mlgi_batch* batch);

/* ---------------------------------------------------------------- autotuner
 * No counterpart in the reference. Picks, per frame shape (G-buffer size,
 * VPL count, precision, shadowing), the interleave and the gather tile:
 *   interleave: the largest of {8, 4} that leaves at least
 *     min_vpls_per_pixel VPLs per interleave class (else 4); it decides
 *     which VPLs a pixel sums, i.e. the image, so it follows a quality rule;
 *   tile: the fastest gather strip width of {8, 16, 32} (workgroup tiles
 *     8x8, 16x4 / 16x16, 32x8), timed on the frame itself through the batch
 *     path; the image does not depend on it.
 * Writes the choice into frame->params (interleave, use_interleaving,
 * tile_w, tile_h). Results are cached per shape in the context; a cached
 * shape is not timed again. Runs the frame (its buffers are overwritten). */
typedef struct mlgi_tune_result {
  int32_t interleave;
  int32_t tile_w, tile_h;
  float gather_ms[3]; /* timed strip widths 8, 16, 32 (0 when cached) */
  int32_t cached;
} mlgi_tune_result;
int mlgi_autotune(mlgi_context* ctx, mlgi_frame_desc* frame, int32_t min_vpls_per_pixel,
                  mlgi_tune_result* out);

/* ---------------------------------------------------------------- serving
 * No counterpart in the reference (an interactive viewer). A GI request
 * server over resident scenes: each scene's BVH, point clouds (one CSR
 * binning per VPL count) and per-resolution G-buffers stay on the device;
 * requests (scene, resolution, VPL count, light pose) arrive over time, are
 * admitted as in-flight slots and a fixed HBM budget free up (their work
 * buffers come from one arena allocated up front), and are batched
 * continuously: every admission round packs all admissible waiting requests
 * into one batch. Queued batches go to `streams` device streams of equal
 * priority in turn, so a new batch runs beside those already on the device
 * instead of behind them. The light pose's RSM is rendered on the device per
 * request. A request's result depends only on the request (the same bits
 * alone or in any mix, on any stream). */
typedef struct mlgi_server mlgi_server;
typedef struct mlgi_server_config {
  size_t hbm_budget;          /* arena bytes the admitted requests' buffers come from */
  int32_t max_inflight;       /* requests on the device at once */
  int32_t max_batch;          /* requests per batch */
  int32_t max_queued_batches; /* batches enqueued on the device at once */
  int32_t rsm_size;           /* RSM side rendered per request */
  int32_t precision;          /* mlgi_precision */
  int32_t min_vpls_per_pixel; /* autotuner interleave rule */
  const int32_t* vpl_counts;  /* VPL counts requests may use (point binnings) */
  int32_t n_vpl_counts;
  const int32_t* resolutions; /* (width, height) pairs requests may use */
  int32_t n_resolutions;
  int32_t streams;            /* device streams the batches are spread over, at most
                                 max_queued_batches; 0 = max_queued_batches */
} mlgi_server_config;
typedef struct mlgi_server_scene {
  const struct mlgi_scene* scene; /* host scene (mlgi_scene.h) */
  int64_t n_points;               /* splat points of the scene */
  uint64_t point_seed;
} mlgi_server_scene;
typedef struct mlgi_request {
  int32_t scene;          /* index into the server's scenes */
  int32_t width, height;  /* one of the configured resolutions (the scene's fixed camera) */
  int32_t n_vpl;          /* one of the configured VPL counts */
  uint64_t light_seed;    /* spot light pose (mlgi_scene_light_pose) */
  double arrival_s;       /* arrival, seconds after mlgi_server_run starts */
  float* host_out;        /* optional: irradiance (h*w*3) copied here on completion */
} mlgi_request;
typedef struct mlgi_server_stats {
  int64_t completed, rejected;
  double wall_s, req_per_s;
  double latency_p50_ms, latency_p99_ms, latency_mean_ms, latency_max_ms;
  int32_t batches, max_inflight_seen;
  double mean_batch;
  size_t peak_bytes;      /* peak arena use */
  int32_t overlapped_batches; /* batches enqueued while another was still on the device */
  int32_t streams;            /* device streams used */
} mlgi_server_stats;
/* Uploads the scenes, builds their point binnings and G-buffers, tunes every
 * (resolution, VPL count) shape. */
int mlgi_server_create(mlgi_context* ctx, const mlgi_server_config* cfg,
                       const mlgi_server_scene* scenes, int32_t n_scenes, mlgi_server** out);
/* Serves the requests (sorted by arrival) in real time and returns when all
 * are completed or rejected (never fit the budget). latencies_ms (optional,
 * n entries) receives each request's arrival-to-completion time (-1 if
 * rejected). */
int mlgi_server_run(mlgi_server* srv, const mlgi_request* reqs, int32_t n, mlgi_server_stats* stats,
                    double* latencies_ms);
void mlgi_server_destroy(mlgi_server* srv);

/* Number of (pixel, VPL-of-its-class) pairs with a positive geometry factor
 * (the lit pairs the gather's algorithmic work is counted on). Synchronous. */
int mlgi_count_lit_pairs(mlgi_context* ctx, const mlgi_gbuffer* gb, const mlgi_vpls* vpls,
                         const mlgi_params* p, uint64_t* out);

/* --------------------------------------------------------------- timing */
/* Enable per-stage HIP event timing (PerfCounter analogue). */
int mlgi_profile_enable(mlgi_context* ctx, int32_t on);
int mlgi_profile_reset(mlgi_context* ctx);
/* Averages over the frames profiled since the last reset. Synchronises. */
int mlgi_profile_read(mlgi_context* ctx, mlgi_stage_times* out);

/* ------------------------------------------------- host-side helpers */
/* The fixed VPL permutation of VPLProcessor.cpp:58-62:
 * std::shuffle(iota(n), std::mt19937(seed)). */
int mlgi_shuffle_indices(int32_t n, uint32_t seed, int32_t* out);
/* ISM grid side: int(pow(2, ceil(log2(ism_count) / 2))) (ImperfectShadowmap.cpp:189). */
int32_t mlgi_ism_indices_1d(int32_t ism_count);

#ifdef __cplusplus
}
#endif

#endif /* MLGI_H */
