/*
 * mlgi scene tooling -- host-side synthetic inputs for the GI hot path.
 *
 * The reference renders its inputs with GL rasterisation (RasterizationStage
 * for the camera G-buffer and the RSM, RasterizationStage.cpp:131-277) and
 * generates splat points with tessellation + geometry shaders
 * (ism.tesc/ism.geom). This header provides seeded, deterministic
 * replacements that produce the same data formats on the host:
 *
 *   mlgi_scene_create          procedural room + random triangles (seeded)
 *   mlgi_scene_render_gbuffer  window depth + encoded face normal per pixel
 *   mlgi_scene_render_rsm      RSM diffuse/normal/depth of a spot light
 *   mlgi_scene_sample_points   area-proportional point set, CSR-binned by
 *                              the reference's Jenkins-hash random()
 *
 * All outputs are host arrays; the caller uploads them. Matrices are
 * column-major float[16] (GL convention), computed in double. The same G-buffer
 * and RSM can also be produced on the device (mlgi_scene_upload below).
 */
#ifndef MLGI_SCENE_H
#define MLGI_SCENE_H

#include <stddef.h>
#include <stdint.h>

#include "mlgi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlgi_scene mlgi_scene;

/* A pinhole camera or spot light: position, look-at target, up vector,
 * vertical field of view (degrees), aspect, clip planes. For a spot light
 * the field of view is the full cone angle. */
typedef struct mlgi_camera {
  float eye[3];
  float center[3];
  float up[3];
  float fovy_deg;
  float aspect;
  float z_near;
  float z_far;
} mlgi_camera;

/* 10x10x10 room (12 triangles, inward normals) plus n_random triangles with
 * one vertex uniform inside the room and two edges of length U[0.1, 0.5] in
 * uniform directions; per-triangle albedo U[0.2, 0.8]^3. */
int mlgi_scene_create(uint64_t seed, int32_t n_random, mlgi_scene** out);
void mlgi_scene_destroy(mlgi_scene* s);
int32_t mlgi_scene_triangle_count(const mlgi_scene* s);
double mlgi_scene_total_area(const mlgi_scene* s);
/* vertices: n*9 floats, normals n*3, albedo n*3 (any may be NULL). */
int mlgi_scene_triangles(const mlgi_scene* s, float* vertices, float* normals, float* albedo);

/* Default camera (fovy 60, near 0.05, far 50) and a seeded spot-light pose
 * (45 degree full cone) near the ceiling aimed at the lower room. */
void mlgi_scene_default_camera(float aspect, mlgi_camera* out);
void mlgi_scene_light_pose(uint64_t seed, mlgi_camera* out);

/* view, projection (glm lookAt/perspective conventions), and the inverse of
 * P*V; any output may be NULL. */
int mlgi_camera_matrices(const mlgi_camera* cam, float* view16, float* proj16, float* vp_inv16);

/* Camera G-buffer by ray casting pixel centres: depth = window depth of the
 * nearest hit (1.0 on a miss), face_normal = n*0.5+0.5 with n facing the
 * camera. threads <= 0 uses the hardware concurrency. */
int mlgi_scene_render_gbuffer(const mlgi_scene* s, const mlgi_camera* cam, int32_t width,
                              int32_t height, int32_t threads, float* depth, float* face_normal);

/* RSM of a spot light: diffuse = albedo inside the cone (0 outside),
 * face_normal = n*0.5+0.5 facing the light, depth = window depth, and
 * inv_bvp16 = inverse(bias * P * V) as VPLProcessor.cpp:76-91 builds it. */
int mlgi_scene_render_rsm(const mlgi_scene* s, const mlgi_camera* light, int32_t width,
                          int32_t height, int32_t threads, float* diffuse, float* face_normal,
                          float* depth, float* inv_bvp16);

/* Area-proportional splat points. Each record is (x, y, z, bits) with bits =
 * packUnorm4x8(n*0.5+0.5, r/25), r = sqrt(A_total / (P*pi)) * sqrt(ism_count)
 * (ism.geom:60-64). Point i goes to bin int(random(vec3(pos)) * n_bins)
 * (random.glsl:62-66); records are grouped by bin (stable in i) and
 * bin_offsets (n_bins+1) delimits the bins. records: n_points*4 floats. */
int mlgi_scene_sample_points(const mlgi_scene* s, int64_t n_points, int32_t n_bins,
                             int32_t ism_count, uint64_t seed, int32_t threads, float* records,
                             uint32_t* bin_offsets);

/* Deferred-shading inputs of a view (model.frag:65-132): albedo (w*h*3, 0
 * for the background) and/or VSM moments (w*h*2: distance d of the hit from
 * the camera's eye and d*d; the reference's screen-space derivative term is
 * left out; background (0, 0)). Rendered from the light, the moments are the
 * light's variance shadow map. Either output may be NULL. */
int mlgi_scene_render_surface(const mlgi_scene* s, const mlgi_camera* cam, int32_t width, int32_t height,
                              int32_t threads, float* albedo, float* moments);

/* ---------------------------------------------------- device producer
 * The scene's BVH on the device, for inputs that change per request or per
 * frame (light poses, views). The kernels run the host renderer's double
 * precision code, so their texels equal mlgi_scene_render_gbuffer /
 * mlgi_scene_render_rsm bit for bit. Output pointers are device buffers;
 * the launches are asynchronous on the context stream. */
typedef struct mlgi_scene_device mlgi_scene_device;
int mlgi_scene_upload(mlgi_context* ctx, const mlgi_scene* s, mlgi_scene_device** out);
void mlgi_scene_device_destroy(mlgi_scene_device* d);
/* depth (w*h), face_normal (w*h*3); vp_inv16 (host, optional) receives
 * inverse(P*V) as mlgi_camera_matrices computes it. */
int mlgi_scene_device_render_gbuffer(mlgi_context* ctx, const mlgi_scene_device* d,
                                     const mlgi_camera* cam, int32_t width, int32_t height,
                                     float* depth, float* face_normal, float* vp_inv16);
/* The point set of mlgi_scene_sample_points generated on the device
 * (ism.tesc:11-23, ism.geom:49-69: the reference emits it per frame from the
 * tessellation and geometry stages). records (n_points*4 floats) and
 * bin_offsets (n_bins+1) are device buffers. Record i and its bin equal the
 * host sampler's; within a bin the records come in no fixed order (the splat
 * does not depend on it), bin_offsets equal the host's. Allocates its work
 * buffers and waits for the stream: scene set-up, not per frame. */
int mlgi_scene_device_sample_points(mlgi_context* ctx, const mlgi_scene_device* d, int64_t n_points,
                                    int32_t n_bins, int32_t ism_count, uint64_t seed, float* records,
                                    uint32_t* bin_offsets);
/* mlgi_scene_render_surface on the device (device buffers, asynchronous). */
int mlgi_scene_device_render_surface(mlgi_context* ctx, const mlgi_scene_device* d, const mlgi_camera* cam,
                                     int32_t width, int32_t height, float* albedo, float* moments);
/* diffuse (w*h*3), face_normal (w*h*3), depth (w*h); inv_bvp16 (host,
 * optional) receives inverse(bias * P * V). */
int mlgi_scene_device_render_rsm(mlgi_context* ctx, const mlgi_scene_device* d,
                                 const mlgi_camera* light, int32_t width, int32_t height,
                                 float* diffuse, float* face_normal, float* depth, float* inv_bvp16);

#ifdef __cplusplus
}
#endif

#endif /* MLGI_SCENE_H */
