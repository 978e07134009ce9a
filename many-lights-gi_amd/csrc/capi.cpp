// C ABI of libmlgi.so: context, device memory, shape validation and the
// stage entry points declared in include/mlgi.h.
//
// Every launch is preceded by host-side checks that the descriptors'
// shapes match what the kernel grids assume, so a kernel never indexes past
// a buffer the caller described.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/mlgi.h"
#include "../../include/mlgi_scene.h"
#include "capi_internal.hpp"
#include "kernels.hpp"
#include "scene_trace.hpp"

using namespace mlgi_capi;

static_assert(MLGI_NOALLOC_MAX_POINTS == mlgi::kScratchMaxPoints, "the ABI's limit is the static area's size");

namespace mlgi_capi {

int fail(mlgi_context* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->error = buf;
  }
  return code;
}

int hip_check(mlgi_context* c, hipError_t e, const char* what) {
  if (c) c->tail = nullptr;  // the call may have enqueued work after the tail event
  if (e == hipSuccess) return MLGI_OK;
  return fail(c, e == hipErrorOutOfMemory ? MLGI_ERR_OUT_OF_MEMORY : MLGI_ERR_HIP, "%s: %s", what,
              hipGetErrorString(e));
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Whether the byte ranges [a, a + na) and [b, b + nb) share a byte.
bool overlaps(const void* a, size_t na, const void* b, size_t nb) {
  const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
  return x < y + nb && y < x + na;
}
int log2i(int v) {
  int k = 0;
  while ((1 << k) < v) ++k;
  return k;
}

hipEvent_t take_event(mlgi_context* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

hipEvent_t tail_event(mlgi_context* c) {
  if (c->tail) return c->tail;
  hipEvent_t e = take_event(c);
  if (!e) return nullptr;
  if (hipEventRecord(e, c->stream) != hipSuccess) {
    c->pool.push_back(e);
    return nullptr;
  }
  c->tail = e;
  return e;
}

// The events of the recorded spans, each once (adjacent spans share one).
static std::vector<hipEvent_t> span_events(const mlgi_context* c) {
  std::vector<hipEvent_t> v;
  for (const auto& s : c->spans) {
    v.push_back(s.a);
    v.push_back(s.b);
  }
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}


mlgi::Mat4 to_mat(const float* m) {
  mlgi::Mat4 r;
  std::memcpy(r.m, m, sizeof r.m);
  return r;
}

int get_shuffle(mlgi_context* c, int n, const int32_t** out) {
  auto it = c->shuffle.find(n);
  if (it != c->shuffle.end()) {
    *out = it->second;
    return MLGI_OK;
  }
  std::vector<int32_t> v(n);
  mlgi_shuffle_indices(n, 1979982u, v.data());
  int32_t* d = nullptr;
  int rc = hip_check(c, hipMalloc(&d, sizeof(int32_t) * n), "hipMalloc(shuffle)");
  if (rc) return rc;
  rc = hip_check(c, hipMemcpy(d, v.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice),
                 "hipMemcpy(shuffle)");
  if (rc) {
    (void)hipFree(d);
    return rc;
  }
  c->shuffle[n] = d;
  *out = d;
  return MLGI_OK;
}

int check_params(mlgi_context* c, const mlgi_params* p) {
  if (!p) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "params is NULL");
  if (!p->use_push_pull)
    return fail(c, MLGI_ERR_UNSUPPORTED, "usePushPull=false (raster ISM path) is not supported");
  // a finite, normal far plane: the splat's and the pull-push's divisions by
  // distances and by zFar are exact (IEEE quotients) for divisors in the
  // range the device division check covers
  if (!(p->z_far >= 0x1p-60f && p->z_far <= 0x1p60f))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "z_far must lie in [2^-60, 2^60]");
  return MLGI_OK;
}

int interleave_of(const mlgi_params* p) { return p->use_interleaving ? p->interleave : 1; }

int check_vpls(mlgi_context* c, const mlgi_vpls* v) {
  if (!v || !v->records) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "vpls is NULL");
  if (v->count <= 0 || v->count > (1 << 20))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "vpl count %d out of range (1..2^20)", v->count);
  return MLGI_OK;
}

// Grows the context's splat work area to at least `bytes`. The stream is
// drained first: a queued splat may still use the old area.
int ensure_splat_work(mlgi_context* c, size_t bytes) {
  if (c->splat_ws_bytes >= bytes) return MLGI_OK;
  int rc = hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (rc) return rc;
  if (c->splat_ws) (void)hipFree(c->splat_ws);
  c->splat_ws = nullptr;
  c->splat_ws_bytes = 0;
  if (hipMalloc(&c->splat_ws, bytes) != hipSuccess) {
    c->splat_ws = nullptr;
    return fail(c, MLGI_ERR_OUT_OF_MEMORY, "hipMalloc(splat work, %zu bytes)", bytes);
  }
  c->splat_ws_bytes = bytes;
  return MLGI_OK;
}

// The VPL range [vplStartIndex, vplEndIndex) of n VPLs and what the ISM
// flags make of it (ism.comp:31-40; final_gathering.comp:40-45): scaleISMs
// lays out one ISM per VPL of the range (VPL v in ISM v - start), else one
// per VPL; pointsOnlyIntoScaledISMs bins the points over the range's VPLs.
struct VplRange {
  int start, end, count;
  int ism_count, ism_offset;
  bool points_only;
};
int vpl_range(mlgi_context* c, const mlgi_params* p, int n, VplRange* r) {
  r->start = p->vpl_start_index;
  r->end = p->vpl_end_index > 0 ? p->vpl_end_index : n;
  r->count = r->end - r->start;
  const bool scale = p->scale_isms != 0;
  r->points_only = p->points_only_into_scaled_isms != 0;
  if ((scale || r->points_only) && !(r->start >= 0 && r->count > 0 && r->end <= n))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "VPL range [%d, %d) must lie in [0, %d) for scaleISMs / "
                "pointsOnlyIntoScaledISMs", r->start, r->end, n);
  r->ism_count = scale ? r->count : n;
  r->ism_offset = scale ? r->start : 0;
  return MLGI_OK;
}

// Atlas geometry: n1d = ism grid side, S = atlas side / n1d.
int atlas_geometry(mlgi_context* c, int side, int n_vpl, int* n1d, int* ism_side) {
  if (!is_pow2(side) || side > 16384)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "atlas side %d must be a power of two <= 16384", side);
  const int g = mlgi_ism_indices_1d(n_vpl);
  if (side % g != 0 || side / g < 2)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "atlas side %d too small for %d ISMs (grid %d)", side,
                n_vpl, g);
  *n1d = g;
  *ism_side = side / g;
  return MLGI_OK;
}

}  // namespace mlgi_capi

extern "C" {

int32_t mlgi_version(void) { return MLGI_VERSION; }

int mlgi_device_count(int32_t* count) {
  if (!count) return MLGI_ERR_INVALID_ARGUMENT;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return MLGI_OK;
}

int mlgi_context_create(int32_t device, mlgi_context** out) {
  if (!out) return MLGI_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return MLGI_ERR_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return MLGI_ERR_HIP;
  auto* c = new mlgi_context();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MLGI_ERR_HIP;
  }
  if (hipMalloc(&c->counter, sizeof(unsigned long long)) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MLGI_ERR_OUT_OF_MEMORY;
  }
  // the device's static splat work area: its address on this device, with
  // the device current (resolved once per device)
  if (mlgi::splat_scratch_init(device, &c->scratch) != hipSuccess) {
    (void)hipFree(c->counter);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MLGI_ERR_HIP;
  }
  *out = c;
  return MLGI_OK;
}

void mlgi_context_destroy(mlgi_context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->shuffle) (void)hipFree(kv.second);
  for (auto e : span_events(c)) (void)hipEventDestroy(e);
  for (auto e : c->pool) (void)hipEventDestroy(e);
  (void)hipFree(c->counter);
  if (c->splat_ws) (void)hipFree(c->splat_ws);
  if (c->order_event) (void)hipEventDestroy(c->order_event);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* mlgi_last_error(const mlgi_context* c) { return c ? c->error.c_str() : ""; }

int mlgi_sync(mlgi_context* c) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  return hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

void* mlgi_stream(mlgi_context* c) { return c ? (void*)c->stream : nullptr; }

int mlgi_context_wait(mlgi_context* c, mlgi_context* after) {
  if (!c || !after) return MLGI_ERR_INVALID_ARGUMENT;
  if (c == after) return MLGI_OK;  // one stream is already in order
  if (c->device != after->device)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "mlgi_context_wait: contexts on devices %d and %d", c->device,
                after->device);
  if (int rc = hip_check(c, hipSetDevice(c->device), "hipSetDevice")) return rc;
  if (!after->order_event) {
    if (int rc = hip_check(c, hipEventCreateWithFlags(&after->order_event, hipEventDisableTiming),
                           "hipEventCreateWithFlags"))
      return rc;
  }
  // A wait applies to the event's latest record at the time of the wait, so
  // one event per context serves every call.
  if (int rc = hip_check(c, hipEventRecord(after->order_event, after->stream), "hipEventRecord")) return rc;
  return hip_check(c, hipStreamWaitEvent(c->stream, after->order_event, 0), "hipStreamWaitEvent");
}

void mlgi_params_default(mlgi_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->light_intensity = 5.0f;
  p->shuffle_lights = 1;
  p->gi_intensity_factor = 3000.0f;
  p->vpl_clamping_value = 0.001f;
  p->vpl_start_index = 0;
  p->vpl_end_index = 0;  // = vpl count
  p->scale_isms = 0;
  p->points_only_into_scaled_isms = 0;
  p->tess_level_factor = 2.0f;
  p->use_push_pull = 1;
  p->enable_shadowing = 1;
  p->use_interleaving = 1;
  p->interleave = 4;
  p->z_far = 50.0f;
  p->atlas_side = 2048;
  p->pull_push_levels = 6;
  p->precision = MLGI_PRECISION_FP32;
  p->tile_w = 0;
  p->tile_h = 0;
  p->gather_fetch = -1;
  p->splat_block = 0;
}

int mlgi_malloc(mlgi_context* c, size_t bytes, void** dptr) {
  if (!c || !dptr) return MLGI_ERR_INVALID_ARGUMENT;
  *dptr = nullptr;
  if (bytes == 0) bytes = 256;
  return hip_check(c, hipMalloc(dptr, bytes), "hipMalloc");
}

int mlgi_free(mlgi_context* c, void* d) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!d) return MLGI_OK;
  (void)hipStreamSynchronize(c->stream);
  return hip_check(c, hipFree(d), "hipFree");
}

int mlgi_memcpy_htod(mlgi_context* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return MLGI_ERR_INVALID_ARGUMENT;
  if (!bytes) return MLGI_OK;
  int rc = hip_check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream),
                     "hipMemcpyAsync(HtoD)");
  if (rc) return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

int mlgi_memcpy_dtoh(mlgi_context* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return MLGI_ERR_INVALID_ARGUMENT;
  if (!bytes) return MLGI_OK;
  int rc = hip_check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream),
                     "hipMemcpyAsync(DtoH)");
  if (rc) return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

int mlgi_memcpy_dtod(mlgi_context* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return MLGI_ERR_INVALID_ARGUMENT;
  if (!bytes) return MLGI_OK;
  return hip_check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream),
                   "hipMemcpyAsync(DtoD)");
}

int mlgi_probe_valu(mlgi_context* c, double* fma_per_s) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!fma_per_s) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "fma_per_s is NULL");
  *fma_per_s = 0.0;
  const int blocks = 2048, iters = 4096;
  float* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = hip_check(c, hipMalloc(&out, blocks * sizeof(float)), "hipMalloc(probe)");
  if (!rc) rc = hip_check(c, hipEventCreate(&e0), "hipEventCreate");
  if (!rc) rc = hip_check(c, hipEventCreate(&e1), "hipEventCreate");
  float best = 0.0f;
  for (int rep = 0; rep < 4 && !rc; ++rep) {  // the first run warms up
    if ((rc = hip_check(c, hipEventRecord(e0, c->stream), "hipEventRecord"))) break;
    if ((rc = hip_check(c, mlgi::launch_valu_probe(c->stream, out, blocks, iters), "valu_probe"))) break;
    if ((rc = hip_check(c, hipEventRecord(e1, c->stream), "hipEventRecord"))) break;
    if ((rc = hip_check(c, hipEventSynchronize(e1), "hipEventSynchronize"))) break;
    float ms = 0.0f;
    if ((rc = hip_check(c, hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime"))) break;
    if (rep > 0 && (best == 0.0f || ms < best)) best = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (out) (void)hipFree(out);
  if (rc) return rc;
  *fma_per_s = best > 0.0f ? (double)mlgi::valu_probe_fmas(blocks, iters) / (best * 1e-3) : 0.0;
  return MLGI_OK;
}

int mlgi_check_division(mlgi_context* c, uint32_t seed, uint64_t* reciprocal_mismatches,
                        uint64_t* quotient_mismatches, uint64_t* quotients_tested) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!reciprocal_mismatches || !quotient_mismatches || !quotients_tested)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "an output of mlgi_check_division is NULL");
  uint32_t* counts = nullptr;
  uint32_t host[2] = {0u, 0u};
  int rc = hip_check(c, hipMalloc(&counts, sizeof(host)), "hipMalloc(division check)");
  if (!rc) rc = hip_check(c, hipMemcpyAsync(counts, host, sizeof(host), hipMemcpyHostToDevice, c->stream),
                          "hipMemcpyAsync(HtoD)");
  if (!rc) rc = hip_check(c, mlgi::launch_division_check(c->stream, counts, seed), "division_check");
  if (!rc) rc = hip_check(c, hipMemcpyAsync(host, counts, sizeof(host), hipMemcpyDeviceToHost, c->stream),
                          "hipMemcpyAsync(DtoH)");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (counts) (void)hipFree(counts);
  if (rc) return rc;
  *reciprocal_mismatches = host[0];
  *quotient_mismatches = host[1];
  *quotients_tested = mlgi::division_check_quotients();
  return MLGI_OK;
}

int mlgi_shuffle_indices(int32_t n, uint32_t seed, int32_t* out) {
  if (n < 0 || (!out && n)) return MLGI_ERR_INVALID_ARGUMENT;
  std::vector<int> v(n);
  std::iota(v.begin(), v.end(), 0);
  std::mt19937 g(seed);
  std::shuffle(v.begin(), v.end(), g);
  for (int i = 0; i < n; ++i) out[i] = v[i];
  return MLGI_OK;
}

int32_t mlgi_ism_indices_1d(int32_t ism_count) {
  int k = 0;
  while ((1LL << (2 * k)) < (long long)ism_count) ++k;
  return 1 << k;
}

int mlgi_vpl_extract(mlgi_context* c, const mlgi_rsm* rsm, const mlgi_params* p, mlgi_vpls* out) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!p) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "params is NULL");
  if (!rsm || !rsm->diffuse || !rsm->face_normal || !rsm->depth)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "rsm buffers are NULL");
  if (rsm->width <= 0 || rsm->height <= 0 || rsm->width > 16384 || rsm->height > 16384)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "rsm size %dx%d out of range", rsm->width, rsm->height);
  int rc = check_vpls(c, out);
  if (rc) return rc;
  const int32_t* sh = nullptr;
  if (p->shuffle_lights && (rc = get_shuffle(c, out->count, &sh))) return rc;
  Scope sc(c, kVpl);
  return hip_check(c,
                   mlgi::launch_vpl_extract(c->stream, rsm->diffuse, rsm->face_normal, rsm->depth,
                                            rsm->width, rsm->height,
                                            to_mat(rsm->inv_biased_view_projection),
                                            p->light_intensity, sh, out->count, out->records),
                   "vpl_extract");
}

}  // extern "C"

namespace mlgi_capi {

// Work area of an ISM-major splat.
enum SplatArea {
  kAreaContext,  // the context's area, grown on demand (or a captured graph's own)
  kAreaStatic    // the device's static area; no allocation, no profiling events
};

int ism_splat(mlgi_context* c, const mlgi_points* pts, const mlgi_vpls* vpls, const mlgi_params* p,
              mlgi_atlas* atlas, SplatArea area) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  int rc = check_params(c, p);
  if (rc || (rc = check_vpls(c, vpls))) return rc;
  if (!pts || !atlas || !atlas->texels)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "points / atlas is NULL");
  if (pts->count < 0 || (pts->count > 0 && (!pts->records || !pts->bin_offsets)))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "points buffers are NULL");
  VplRange r;
  if ((rc = vpl_range(c, p, vpls->count, &r))) return rc;
  const int bins = r.points_only ? r.count : vpls->count;  // sampledVplCount
  if (pts->count > 0 && pts->n_bins != bins)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "point bins (%d) must equal the %s (%d)", pts->n_bins,
                r.points_only ? "VPL range's count" : "VPL count", bins);
  if (p->splat_block != 0 && p->splat_block != 64 && p->splat_block != 128 && p->splat_block != 256)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "splat_block %d not in {0, 64, 128, 256}", p->splat_block);
  int n1d, s;
  if ((rc = atlas_geometry(c, atlas->side, r.ism_count, &n1d, &s))) return rc;
  if (atlas->ism_side == 0) atlas->ism_side = s;
  if (atlas->ism_side != s)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "atlas ism_side %d != %d for %d ISMs in a %d atlas",
                atlas->ism_side, s, r.ism_count, atlas->side);
  const size_t texels = (size_t)atlas->side * atlas->side;
  // the ISM-major form assumes one ISM per VPL with VPL v's points in the
  // bins v - i: the VPL-range flags take the point-major form
  const bool ranged = p->scale_isms || p->points_only_into_scaled_isms;
  if (p->splat_block == 0 && !ranged && mlgi::ism_major_supported(s) && ((uintptr_t)atlas->texels & 15u) == 0) {
    // ISM-major form: writes every texel, no clear
    const uint32_t np = pts->count > 0 ? (uint32_t)pts->count : 0u;
    const size_t need = mlgi::splat_work_bytes(np);
    void* ws = nullptr;
    if (area == kAreaStatic) {
      ws = c->scratch;  // sized for kScratchMaxPoints (checked by the caller)
    } else if (c->capture_ws) {
      if (c->capture_ws_bytes < need)
        return fail(c, MLGI_ERR_INVALID_ARGUMENT, "captured splat needs %zu work bytes, the graph holds %zu", need,
                    c->capture_ws_bytes);
      ws = c->capture_ws;
    } else {
      if ((rc = ensure_splat_work(c, need))) return rc;
      ws = c->splat_ws;
      c->splat_ws_points = np;
    }
    Scope sc(c, area == kAreaStatic ? -1 : kSplat);
    return hip_check(c,
                     mlgi::launch_ism_splat_tiled(c->stream, pts->records, pts->bin_offsets,
                                                  pts->count > 0 ? pts->n_bins : 0, pts->count, vpls->records,
                                                  vpls->count, n1d, s, atlas->side, p->z_far, atlas->texels,
                                                  mlgi::splat_work(ws, np)),
                     "ism_splat_tiled");
  }
  {
    Scope sc(c, area == kAreaStatic ? -1 : kClear);
    if ((rc = hip_check(c, mlgi::launch_atlas_clear(c->stream, atlas->texels, texels), "atlas_clear")))
      return rc;
  }
  if (pts->count == 0) return MLGI_OK;
  Scope sc(c, area == kAreaStatic ? -1 : kSplat);
  return hip_check(c,
                   mlgi::launch_ism_splat(c->stream, pts->records, pts->bin_offsets, pts->n_bins,
                                          pts->count, vpls->records, vpls->count, n1d, atlas->side,
                                          p->z_far, atlas->texels,
                                          p->splat_block > 0 ? p->splat_block : 256, r.start, r.count,
                                          r.ism_offset, r.points_only ? 1 : 0),
                   "ism_splat");
}

}  // namespace mlgi_capi

extern "C" {

int mlgi_ism_splat(mlgi_context* c, const mlgi_points* pts, const mlgi_vpls* vpls, const mlgi_params* p,
                   mlgi_atlas* atlas) {
  return ism_splat(c, pts, vpls, p, atlas, kAreaContext);
}

int mlgi_ism_splat_noalloc(mlgi_context* c, const mlgi_points* pts, const mlgi_vpls* vpls, const mlgi_params* p,
                           mlgi_atlas* atlas) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (pts && (int64_t)pts->count > (int64_t)MLGI_NOALLOC_MAX_POINTS)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT,
                "mlgi_ism_splat_noalloc: %d points exceed the static work area's maximum of %d", pts->count,
                (int)MLGI_NOALLOC_MAX_POINTS);
  if (!c->scratch) return fail(c, MLGI_ERR_HIP, "mlgi_ism_splat_noalloc: no static work area on this device");
  return ism_splat(c, pts, vpls, p, atlas, kAreaStatic);
}

int mlgi_deferred_composite(mlgi_context* c, const mlgi_composite_inputs* in, const mlgi_composite_params* p,
                            mlgi_image* out) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!in || !p || !out || !out->rgb) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!in->depth || !in->normal || !in->diffuse || !in->gi || !in->vsm)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "composite input buffer is NULL");
  if (in->width <= 0 || in->height <= 0 || in->width > 16384 || in->height > 16384 || in->vsm_width <= 0 ||
      in->vsm_height <= 0 || in->vsm_width > 16384 || in->vsm_height > 16384)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "composite sizes out of range");
  if (out->width != in->width || out->height != in->height)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "output %dx%d != input %dx%d", out->width, out->height, in->width,
                in->height);
  if (((uintptr_t)in->vsm & 7u) != 0) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "vsm must be 8-byte aligned");
  mlgi::CompositeArgs a;
  a.depth = in->depth;
  a.normal = in->normal;
  a.diffuse = in->diffuse;
  a.specular = in->specular;
  a.gi = in->gi;
  a.occlusion = in->occlusion;
  a.w = in->width;
  a.h = in->height;
  a.vsm = in->vsm;
  a.vw = in->vsm_width;
  a.vh = in->vsm_height;
  std::memcpy(a.vp_inv, p->view_projection_inverse, sizeof a.vp_inv);
  std::memcpy(a.blvp, p->biased_light_view_projection, sizeof a.blvp);
  std::memcpy(a.eye, p->camera_eye, sizeof a.eye);
  std::memcpy(a.light_pos, p->light_position, sizeof a.light_pos);
  std::memcpy(a.inv_light_dir, p->inverse_light_direction, sizeof a.inv_light_dir);
  a.light_intensity = p->light_intensity;
  a.exposure = p->exposure;
  a.out = out->rgb;
  return hip_check(c, mlgi::launch_composite(c->stream, a), "deferred_composite");
}

int mlgi_splat_spill_stats(mlgi_context* c, uint32_t* spills, int32_t* overflowed) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!spills || !overflowed) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "NULL argument");
  *spills = 0;
  *overflowed = 0;
  if (!c->splat_ws) return MLGI_OK;
  uint32_t st[2] = {0, 0};
  int rc = hip_check(c, hipMemcpyAsync(st, c->splat_ws, sizeof st, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (rc) return rc;
  *spills = st[0];
  *overflowed = st[1] ? 1 : 0;
  return MLGI_OK;
}

int mlgi_splat_count(mlgi_context* c, uint64_t* splats) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!splats) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "splats is NULL");
  *splats = 0;
  if (!c->splat_ws || c->splat_ws_points < 0)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "no ISM-major splat has run through mlgi_ism_splat on this context");
  const uint32_t np = (uint32_t)c->splat_ws_points;
  int rc = hip_check(c, hipMemsetAsync(c->counter, 0, sizeof(unsigned long long), c->stream), "hipMemsetAsync");
  if (!rc)
    rc = hip_check(c, mlgi::launch_count_splats(c->stream, mlgi::splat_work(c->splat_ws, np), np, c->counter),
                   "count_splats");
  unsigned long long v = 0;
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(&v, c->counter, sizeof v, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (!rc) *splats = v;
  return rc;
}

int mlgi_ism_pullpush(mlgi_context* c, const mlgi_atlas* atlas, const mlgi_params* p,
                      mlgi_ism_depth* out) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  int rc = check_params(c, p);
  if (rc) return rc;
  if (!atlas || !atlas->texels || !out || !out->texels)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "atlas / output is NULL");
  if (out->side != atlas->side)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "output side %d != atlas side %d", out->side, atlas->side);
  const int s = atlas->ism_side;
  if (!is_pow2(atlas->side) || atlas->side > 16384)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "atlas side %d must be a power of two <= 16384", atlas->side);
  if (s != 8 && s != 16 && s != 32 && s != 64)
    return fail(c, MLGI_ERR_UNSUPPORTED, "ISM side %d not in {8,16,32,64}", s);
  if (atlas->side % s != 0)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "ISM side %d does not divide atlas side %d", s, atlas->side);
  if (p->pull_push_levels < 1 || p->pull_push_levels > log2i(s))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "pull-push levels %d out of range 1..%d",
                p->pull_push_levels, log2i(s));
  // the alignment every entry point states for these buffers (the splat's
  // tile pass writes the atlas in 16-byte units; the pull-push reads it in 8-
  // and writes the ISM in 2-byte units)
  if (((uintptr_t)atlas->texels & 15u) != 0 || ((uintptr_t)out->texels & 7u) != 0)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "atlas must be 16-byte and ISM 8-byte aligned");
  // the level-0 push writes the result while other workgroups still read the atlas
  const size_t texels = (size_t)atlas->side * (size_t)atlas->side;
  if (overlaps(atlas->texels, texels * sizeof(uint32_t), out->texels, texels * sizeof(uint16_t)))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "the pull-push output overlaps the atlas");
  out->ism_side = s;
  Scope sc(c, kPullPush);
  return hip_check(c,
                   mlgi::launch_pullpush(c->stream, atlas->texels, atlas->side, s,
                                         p->pull_push_levels, p->z_far, out->texels),
                   "pullpush");
}

int mlgi_ism_process(mlgi_context* c, const mlgi_points* pts, const mlgi_vpls* vpls,
                     const mlgi_params* p, mlgi_atlas* atlas, mlgi_ism_depth* out) {
  int rc = mlgi_ism_splat(c, pts, vpls, p, atlas);
  if (rc) return rc;
  return mlgi_ism_pullpush(c, atlas, p, out);
}

}  // extern "C"

namespace mlgi_capi {

int build_gather_args(mlgi_context* c, const mlgi_gbuffer* gb, const mlgi_vpls* vpls,
                      const mlgi_ism_depth* ism, const mlgi_params* p, mlgi::GatherArgs* a) {
  int rc = check_params(c, p);
  if (rc || (rc = check_vpls(c, vpls))) return rc;
  if (!gb || !gb->depth || !gb->face_normal)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "gbuffer buffers are NULL");
  if (gb->width <= 0 || gb->height <= 0 || gb->width > 16384 || gb->height > 16384)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "gbuffer size %dx%d out of range", gb->width, gb->height);
  const int il = interleave_of(p);
  if (!is_pow2(il) || il > 16)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "interleave %d must be a power of two <= 16", il);
  const int classes = il * il;
  if (vpls->count % classes != 0)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "vpl count %d not divisible by interleave classes %d",
                vpls->count, classes);
  const int vpl_end = p->vpl_end_index > 0 ? p->vpl_end_index : vpls->count;
  if (vpl_end - p->vpl_start_index <= 0)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "vplEndIndex - vplStartIndex must be > 0");
  std::memset(a, 0, sizeof *a);
  if (p->enable_shadowing) {
    if (!ism || !ism->texels) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "ism is NULL");
    VplRange r;
    if ((rc = vpl_range(c, p, vpls->count, &r))) return rc;
    int n1d, s;
    if ((rc = atlas_geometry(c, ism->side, r.ism_count, &n1d, &s))) return rc;
    if (ism->ism_side != s)
      return fail(c, MLGI_ERR_INVALID_ARGUMENT, "ism side %d != %d for %d ISMs in a %d atlas",
                  ism->ism_side, s, r.ism_count, ism->side);
    a->ism_offset = r.ism_offset;
    a->ism = ism->texels;
    a->atlas_side = ism->side;
    a->ism_side = s;
    a->n1d = n1d;
    a->n1d_log2 = log2i(n1d);
  } else {
    a->atlas_side = 1;
    a->ism_side = 1;
    a->n1d = 1;
  }
  a->depth = gb->depth;
  a->normal = gb->face_normal;
  a->width = gb->width;
  a->height = gb->height;
  a->vp_inv = to_mat(gb->view_projection_inverse);
  a->vpl = vpls->records;
  a->n_vpl = vpls->count;
  a->per_class = vpls->count / classes;
  a->interleave = il;
  a->interleave_log2 = log2i(il);
  a->z_far = p->z_far;
  // smoothstep(1 - bias, 1, 1 - (d/zFar - occ)) = smoothstep of t = 1 + (occ - d/zFar) / bias
  const double bias = (double)0.02f;
  a->shadow_k1 = (float)(1.0 / (bias * (double)p->z_far));
  a->shadow_k2 = (float)(1.0 / (bias * 65535.0));
  a->clamp_value = p->vpl_clamping_value;
  a->out_scale = p->gi_intensity_factor / (float)(vpl_end - p->vpl_start_index) * (float)classes;
  a->shadowing = p->enable_shadowing ? 1 : 0;
  a->cull = p->disable_culling ? 0 : 1;
  return MLGI_OK;
}

void pick_tile(const mlgi_params* p, int* tw, int* th) {
  *tw = p->tile_w > 0 ? p->tile_w : 8;
  *th = p->tile_h > 0 ? p->tile_h : 8;
}

}  // namespace mlgi_capi

int check_precision(mlgi_context* c, const mlgi_params* p) {
  if (p->precision != MLGI_PRECISION_FP32 && p->precision != MLGI_PRECISION_MIXED)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "unknown precision %d", p->precision);
  return MLGI_OK;
}

extern "C" {

int mlgi_final_gather(mlgi_context* c, const mlgi_gbuffer* gb, const mlgi_vpls* vpls,
                      const mlgi_ism_depth* ism, const mlgi_params* p, mlgi_image* out) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  mlgi::GatherArgs a;
  int rc = build_gather_args(c, gb, vpls, ism, p, &a);
  if (rc) return rc;
  if (!out || !out->rgb || out->width != gb->width || out->height != gb->height)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "output image must match the gbuffer size");
  if ((rc = check_precision(c, p))) return rc;
  int tw, th;
  pick_tile(p, &tw, &th);
  if (!((tw == 8 && th == 8) || (tw == 16 && th == 16) || (tw == 32 && th == 8) || (tw == 16 && th == 4)))
    return fail(c, MLGI_ERR_UNSUPPORTED, "tile %dx%d not in {8x8, 16x16, 32x8, 16x4}", tw, th);
  const int fetch = p->gather_fetch >= 0 ? p->gather_fetch : 0;
  if (fetch > 1) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "gather_fetch %d not in 0..1", fetch);
  a.out = out->rgb;
  Scope sc(c, kGather);
  return hip_check(c, mlgi::launch_final_gather(c->stream, a, p->precision, tw, th, fetch), "final_gather");
}

int mlgi_gi_blur(mlgi_context* c, const mlgi_gbuffer* gb, const float* projection16, const mlgi_image* gi,
                 mlgi_image* out) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!gb || !gb->depth || !gb->face_normal || !projection16 || !gi || !gi->rgb || !out || !out->rgb)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "NULL argument");
  if (gb->width <= 0 || gb->height <= 0 || gb->width > 16384 || gb->height > 16384)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "gbuffer size %dx%d out of range", gb->width, gb->height);
  if (gi->width != gb->width || gi->height != gb->height || out->width != gb->width || out->height != gb->height)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "images must match the gbuffer size");
  if (gi->rgb == out->rgb) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "the blur cannot run in place");
  Scope sc(c, kBlur);
  return hip_check(c, mlgi::launch_gi_blur(c->stream, gi->rgb, gb->face_normal, gb->depth, gb->width, gb->height,
                                           projection16, out->rgb),
                   "gi_blur");
}

int mlgi_frame(mlgi_context* c, const mlgi_rsm* rsm, const mlgi_points* pts,
               const mlgi_gbuffer* gb, const mlgi_params* p, mlgi_vpls* vpls, mlgi_atlas* atlas,
               mlgi_ism_depth* ism, mlgi_image* out) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  Scope sc(c, kFrame);
  int rc = mlgi_vpl_extract(c, rsm, p, vpls);
  if (rc) return rc;
  if ((rc = mlgi_ism_process(c, pts, vpls, p, atlas, ism))) return rc;
  if ((rc = mlgi_final_gather(c, gb, vpls, ism, p, out))) return rc;
  if (c->profile) c->frames++;
  return MLGI_OK;
}

}  // extern "C"

struct mlgi_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void* splat_ws = nullptr;  // a captured frame's own splat work area
};

extern "C" {

int mlgi_graph_capture_frame(mlgi_context* c, const mlgi_rsm* rsm, const mlgi_points* pts,
                             const mlgi_gbuffer* gb, const mlgi_params* p, mlgi_vpls* vpls,
                             mlgi_atlas* atlas, mlgi_ism_depth* ism, mlgi_image* out,
                             mlgi_graph** graph) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!graph) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "graph is NULL");
  *graph = nullptr;
  int rc = check_vpls(c, vpls);
  if (rc) return rc;
  if (!p) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "params is NULL");
  if (!pts || pts->count < 0) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "points is NULL or negative");
  // everything that allocates or copies happens before the capture: the
  // shuffle table, and the graph's own splat work area (the context's
  // shared one may be reallocated by a later eager splat, so a graph never
  // refers to it)
  const int32_t* sh = nullptr;
  if (p->shuffle_lights && (rc = get_shuffle(c, vpls->count, &sh))) return rc;
  auto* out_graph = new mlgi_graph();
  const size_t ws_bytes = mlgi::splat_work_bytes((uint32_t)pts->count);
  if ((rc = hip_check(c, hipMalloc(&out_graph->splat_ws, ws_bytes), "hipMalloc(graph splat work)"))) {
    mlgi_graph_destroy(out_graph);
    return rc;
  }
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize"))) {
    mlgi_graph_destroy(out_graph);
    return rc;
  }
  const bool prof = c->profile;
  c->profile = false;  // no event nodes in the graph
  rc = hip_check(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal),
                 "hipStreamBeginCapture");
  if (rc) {
    c->profile = prof;
    mlgi_graph_destroy(out_graph);
    return rc;
  }
  c->capture_ws = out_graph->splat_ws;
  c->capture_ws_bytes = ws_bytes;
  const int frame_rc = mlgi_frame(c, rsm, pts, gb, p, vpls, atlas, ism, out);
  c->capture_ws = nullptr;
  c->capture_ws_bytes = 0;
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(c->stream, &g);
  c->profile = prof;
  out_graph->graph = g;
  if (frame_rc) {
    mlgi_graph_destroy(out_graph);
    return frame_rc;
  }
  if ((rc = hip_check(c, end, "hipStreamEndCapture")) ||
      (rc = hip_check(c, hipGraphInstantiate(&out_graph->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate"))) {
    mlgi_graph_destroy(out_graph);
    return rc;
  }
  *graph = out_graph;
  return MLGI_OK;
}

int mlgi_graph_launch(mlgi_context* c, mlgi_graph* graph) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!graph || !graph->exec) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "graph is NULL");
  Scope sc(c, kFrame);
  int rc = hip_check(c, hipGraphLaunch(graph->exec, c->stream), "hipGraphLaunch");
  if (!rc && c->profile) c->frames++;
  return rc;
}

void mlgi_graph_destroy(mlgi_graph* graph) {
  if (!graph) return;
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  // hipFree waits for the device's queued work, so no replay still uses it
  if (graph->splat_ws) (void)hipFree(graph->splat_ws);
  delete graph;
}

int mlgi_count_lit_pairs(mlgi_context* c, const mlgi_gbuffer* gb, const mlgi_vpls* vpls,
                         const mlgi_params* p, uint64_t* out) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!out) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "out is NULL");
  mlgi_params q = *p;
  q.enable_shadowing = 0;
  mlgi::GatherArgs a;
  int rc = build_gather_args(c, gb, vpls, nullptr, &q, &a);
  if (rc) return rc;
  if ((rc = hip_check(c, hipMemsetAsync(c->counter, 0, sizeof(unsigned long long), c->stream),
                      "hipMemsetAsync")))
    return rc;
  if ((rc = hip_check(c, mlgi::launch_count_lit_pairs(c->stream, a, c->counter), "count_lit_pairs")))
    return rc;
  unsigned long long v = 0;
  if ((rc = hip_check(c, hipMemcpyAsync(&v, c->counter, sizeof v, hipMemcpyDeviceToHost, c->stream),
                      "hipMemcpyAsync")))
    return rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize"))) return rc;
  *out = v;
  return MLGI_OK;
}

int mlgi_profile_enable(mlgi_context* c, int32_t on) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  c->profile = on != 0;
  return MLGI_OK;
}

int mlgi_profile_reset(mlgi_context* c) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  (void)hipStreamSynchronize(c->stream);
  for (auto e : span_events(c)) c->pool.push_back(e);
  c->spans.clear();
  c->tail = nullptr;
  c->frames = 0;
  return MLGI_OK;
}

int mlgi_profile_read(mlgi_context* c, mlgi_stage_times* out) {
  if (!c || !out) return MLGI_ERR_INVALID_ARGUMENT;
  int rc = hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (rc) return rc;
  double sum[kStages] = {0};
  int cnt[kStages] = {0};
  for (auto& s : c->spans) {
    float ms = 0.0f;
    if ((rc = hip_check(c, hipEventElapsedTime(&ms, s.a, s.b), "hipEventElapsedTime"))) return rc;
    sum[s.stage] += ms;
    cnt[s.stage]++;
  }
  // averages per call of each stage (per frame when driven by mlgi_frame)
  auto avg = [&](int st) { return cnt[st] ? (float)(sum[st] / cnt[st]) : 0.0f; };
  out->vpl_ms = avg(kVpl);
  out->ism_clear_ms = avg(kClear);
  out->ism_splat_ms = avg(kSplat);
  out->pullpush_ms = avg(kPullPush);
  out->final_gather_ms = avg(kGather);
  out->blur_ms = avg(kBlur);
  out->frame_ms = avg(kFrame);
  out->frames = c->frames;
  return MLGI_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batches

struct mlgi_batch {
  mlgi_context* ctx = nullptr;
  mlgi::BatchFrame* dev = nullptr;  // device copy of the frame table
  char* work = nullptr;             // the frames' splat work areas
  mlgi_capi::BatchPlan plan;
};

namespace mlgi_capi {

// Validates one frame of a batch and fills its table row (first_* unset).
int batch_frame(mlgi_context* c, const mlgi_frame_desc& d, mlgi::BatchFrame* f) {
  const mlgi_params* p = &d.params;
  int rc = check_params(c, p);
  if (!rc && (p->scale_isms || p->points_only_into_scaled_isms))
    return fail(c, MLGI_ERR_UNSUPPORTED,
                "scaleISMs / pointsOnlyIntoScaledISMs run through the single-frame entry points, not batches");
  if (rc || (rc = check_vpls(c, &d.vpls))) return rc;
  const mlgi_rsm& rsm = d.rsm;
  if (!rsm.diffuse || !rsm.face_normal || !rsm.depth)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "rsm buffers are NULL");
  if (rsm.width <= 0 || rsm.height <= 0 || rsm.width > 16384 || rsm.height > 16384)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "rsm size %dx%d out of range", rsm.width, rsm.height);
  const mlgi_points& pts = d.points;
  if (pts.count < 0 || (pts.count > 0 && (!pts.records || !pts.bin_offsets)))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "points buffers are NULL");
  if (pts.count > 0 && pts.n_bins != d.vpls.count)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "point bins (%d) must equal the VPL count (%d)", pts.n_bins,
                d.vpls.count);
  if (!d.atlas.texels || !d.ism.texels) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "atlas / ism is NULL");
  // the batched tile pass writes the atlas in 16-byte units (and the
  // alignment of mlgi_ism_pullpush holds for the ISM)
  if (((uintptr_t)d.atlas.texels & 15u) != 0 || ((uintptr_t)d.ism.texels & 7u) != 0)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "atlas must be 16-byte and ISM 8-byte aligned");
  // the pull-push writes the ISM while other workgroups still read the atlas
  if (d.atlas.side > 0 && d.ism.side > 0 &&
      overlaps(d.atlas.texels, (size_t)d.atlas.side * d.atlas.side * sizeof(uint32_t), d.ism.texels,
               (size_t)d.ism.side * d.ism.side * sizeof(uint16_t)))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "the ISM overlaps the atlas");
  int n1d, s;
  if ((rc = atlas_geometry(c, d.atlas.side, d.vpls.count, &n1d, &s))) return rc;
  if ((d.atlas.ism_side != 0 && d.atlas.ism_side != s) || d.ism.side != d.atlas.side ||
      (d.ism.ism_side != 0 && d.ism.ism_side != s))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "atlas / ism geometry does not match %d VPLs in a %d atlas",
                d.vpls.count, d.atlas.side);
  if (s != 8 && s != 16 && s != 32 && s != 64)
    return fail(c, MLGI_ERR_UNSUPPORTED, "ISM side %d not in {8,16,32,64}", s);
  if (p->pull_push_levels < 1 || p->pull_push_levels > log2i(s))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "pull-push levels %d out of range 1..%d", p->pull_push_levels,
                log2i(s));
  mlgi_ism_depth ism = d.ism;
  ism.ism_side = s;
  mlgi::GatherArgs a;
  if ((rc = build_gather_args(c, &d.gbuffer, &d.vpls, &ism, p, &a))) return rc;
  if (!d.image.rgb || d.image.width != d.gbuffer.width || d.image.height != d.gbuffer.height)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "output image must match the gbuffer size");
  if ((rc = check_precision(c, p))) return rc;
  const int tw = p->tile_w > 0 ? p->tile_w : 8;
  if (tw != 8 && tw != 16 && tw != 32)
    return fail(c, MLGI_ERR_UNSUPPORTED, "batched gather strip width %d not in {8, 16, 32}", tw);
  const int32_t* sh = nullptr;
  if (p->shuffle_lights && (rc = get_shuffle(c, d.vpls.count, &sh))) return rc;

  std::memset(f, 0, sizeof *f);
  f->rsm_diffuse = rsm.diffuse;
  f->rsm_normal = rsm.face_normal;
  f->rsm_depth = rsm.depth;
  f->rsm_w = rsm.width;
  f->rsm_h = rsm.height;
  f->inv_bvp = to_mat(rsm.inv_biased_view_projection);
  f->light_intensity = p->light_intensity;
  f->shuffle = sh;
  uint32_t s1d = 0;
  while ((uint64_t)s1d * s1d < (uint64_t)d.vpls.count) ++s1d;  // uint(ceil(sqrt(totalVplCount)))
  f->axis_x = s1d;
  f->axis_y = (uint32_t)d.vpls.count / s1d;
  f->points = pts.count > 0 ? pts.records : nullptr;
  f->bin_offsets = pts.count > 0 ? pts.bin_offsets : nullptr;
  f->n_points = (uint32_t)pts.count;
  f->atlas = d.atlas.texels;
  f->atlas_side = d.atlas.side;
  f->ism_side = s;
  f->n1d = n1d;
  f->levels = p->pull_push_levels;
  f->z_far = p->z_far;
  mlgi::pullpush_empty_texel(s, p->z_far, f->pp_e1);
  a.out = d.image.rgb;
  a.ism = d.ism.texels;  // the pull-push writes it even when the gather does not shadow
  f->g = a;
  f->wave_w_log2 = log2i(tw);
  const int cw = (a.width + a.interleave - 1) / a.interleave;
  f->strips_x = (cw + tw - 1) / tw;
  return MLGI_OK;
}

int batch_plan(mlgi_context* c, const mlgi_frame_desc* frames, int count, std::vector<mlgi::BatchFrame>* table,
               BatchPlan* plan) {
  if (!frames || count <= 0 || count > (1 << 16))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "batch of %d frames", count);
  std::vector<mlgi::BatchFrame> rows(count);
  std::vector<int> key(count);
  for (int i = 0; i < count; ++i) {
    int rc = batch_frame(c, frames[i], &rows[i]);
    if (rc) {
      std::string m = c->error;
      return fail(c, rc, "frame %d: %s", i, m.c_str());
    }
    key[i] = (frames[i].params.precision << 16) | (rows[i].ism_side << 8) | rows[i].levels;
  }
  // group by (precision, ISM side, levels): contiguous runs per launch
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key[x] < key[y]; });
  *plan = BatchPlan{};
  plan->count = count;
  plan->row_of.assign(count, 0);
  for (int r = 0; r < count; ++r) plan->row_of[order[r]] = r;
  table->assign(count, mlgi::BatchFrame{});
  uint64_t tv = 0, tt = 0, tb = 0;
  size_t work = 0;
  for (int r = 0; r < count; ++r) {
    mlgi::BatchFrame f = rows[order[r]];
    const mlgi::BatchCounts n = mlgi::batch_counts(f);
    f.first_vpl = (uint32_t)tv;
    f.first_tile = (uint32_t)tt;
    f.first_bin = (uint32_t)tb;
    tv += n.vpl;
    tt += n.tiles;
    tb += n.bins;
    work += mlgi::splat_work_bytes(f.n_points);
    plan->max_ism_side = std::max(plan->max_ism_side, f.ism_side);
    const int k = key[order[r]];
    const int prec = k >> 16;
    if (r == 0 || key[order[r - 1]] != k) plan->pullpush.push_back({r, 0, 0, f.ism_side, f.levels});
    auto& pp = plan->pullpush.back();
    f.first_pp = pp.total;
    pp.total += n.pp;
    pp.count++;
    if (r == 0 || (key[order[r - 1]] >> 16) != prec) plan->gather.push_back({r, 0, 0, prec, 0});
    auto& fg = plan->gather.back();
    f.first_wave = fg.total;
    fg.total += n.waves;
    fg.count++;
    (*table)[r] = f;
  }
  if (tv > 0xFFFFFFFFull || tt > 0xFFFFFFFFull || tb > 0xFFFFFFFFull)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "batch too large");
  plan->total_vpl = (uint32_t)tv;
  plan->total_tiles = (uint32_t)tt;
  plan->total_bins = (uint32_t)tb;
  plan->work_bytes = work;
  return MLGI_OK;
}

void batch_attach_work(std::vector<mlgi::BatchFrame>* table, char* base) {
  size_t off = 0;
  for (auto& f : *table) {
    const mlgi::SplatWork w = mlgi::splat_work(base + off, f.n_points);
    f.mask = w.mask;
    f.spill_state = w.state;
    f.spill = w.spill;
    off += mlgi::splat_work_bytes(f.n_points);
  }
}

int batch_enqueue(mlgi_context* c, const mlgi::BatchFrame* dev, const BatchPlan& plan) {
  Scope frame(c, kFrame);
  int rc;
  {
    Scope sc(c, kVpl);
    if ((rc = hip_check(c, mlgi::launch_batch_vpl(c->stream, dev, plan.count, plan.total_vpl), "batch_vpl")))
      return rc;
  }
  {
    Scope sc(c, kSplat);  // ISM-major: mask, tiles (which clear the atlases), spills
    if ((rc = hip_check(c, mlgi::launch_batch_mask(c->stream, dev, plan.count, plan.total_bins), "batch_mask")) ||
        (rc = hip_check(c, mlgi::launch_batch_tiles(c->stream, dev, plan.count, plan.total_tiles, plan.max_ism_side),
                          "batch_tiles")) ||
        (rc = hip_check(c, mlgi::launch_batch_spill(c->stream, dev, plan.count, plan.total_bins), "batch_spill")))
      return rc;
  }
  {
    Scope sc(c, kPullPush);
    for (const auto& r : plan.pullpush)
      if ((rc = hip_check(c, mlgi::launch_batch_pullpush(c->stream, dev + r.begin, r.count, r.total, r.a, r.b),
                          "batch_pullpush")))
        return rc;
  }
  {
    Scope sc(c, kGather);
    for (const auto& r : plan.gather)
      if ((rc = hip_check(c, mlgi::launch_batch_gather(c->stream, dev + r.begin, r.count, r.total, r.a),
                          "batch_gather")))
        return rc;
  }
  if (c->profile) c->frames++;
  return MLGI_OK;
}

}  // namespace mlgi_capi

mlgi::BatchCounts mlgi::batch_counts(const mlgi::BatchFrame& f) {
  BatchCounts n;
  n.vpl = (uint32_t)f.g.n_vpl;
  n.tiles = (uint32_t)(f.n1d * f.n1d);
  n.bins = f.n_points > 0 ? (uint32_t)f.g.n_vpl : 0u;
  n.pp = (uint32_t)(f.n1d * f.n1d / mlgi::pullpush_group(f.ism_side, f.n1d * f.n1d));
  const int th = 64 >> f.wave_w_log2;
  const int ch = (f.g.height + f.g.interleave - 1) / f.g.interleave;
  const uint32_t strips_y = (uint32_t)((ch + th - 1) / th);
  n.waves = (uint32_t)f.strips_x * strips_y * (uint32_t)(f.g.interleave * f.g.interleave);
  return n;
}

extern "C" {

int mlgi_batch_create(mlgi_context* c, const mlgi_frame_desc* frames, int32_t count, mlgi_batch** out) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!out) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  std::vector<mlgi::BatchFrame> table;
  auto* b = new mlgi_batch();
  b->ctx = c;
  int rc = batch_plan(c, frames, count, &table, &b->plan);
  if (!rc) rc = hip_check(c, hipMalloc(&b->work, b->plan.work_bytes), "hipMalloc(batch work)");
  if (!rc) batch_attach_work(&table, b->work);
  if (!rc) rc = hip_check(c, hipMalloc(&b->dev, sizeof(mlgi::BatchFrame) * count), "hipMalloc(batch)");
  if (!rc)
    rc = hip_check(c, hipMemcpy(b->dev, table.data(), sizeof(mlgi::BatchFrame) * count, hipMemcpyHostToDevice),
                   "hipMemcpy(batch)");
  if (rc) {
    mlgi_batch_destroy(b);
    return rc;
  }
  *out = b;
  return MLGI_OK;
}

int mlgi_batch_run(mlgi_context* c, mlgi_batch* b) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!b || !b->dev || b->ctx != c) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "batch is NULL or foreign");
  return batch_enqueue(c, b->dev, b->plan);
}

int mlgi_batch_update_light(mlgi_context* c, mlgi_batch* b, int32_t frame, const float* inv_bvp16) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!b || !b->dev || b->ctx != c || !inv_bvp16) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "NULL argument");
  if (frame < 0 || frame >= b->plan.count) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "frame %d out of range", frame);
  mlgi::BatchFrame* row = b->dev + b->plan.row_of[frame];
  return hip_check(c, hipMemcpyAsync(&row->inv_bvp, inv_bvp16, sizeof(mlgi::Mat4), hipMemcpyHostToDevice, c->stream),
                   "hipMemcpyAsync(light)");
}

int mlgi_batch_capture(mlgi_context* c, mlgi_batch* b, mlgi_graph** graph) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!graph) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "graph is NULL");
  *graph = nullptr;
  if (!b || !b->dev || b->ctx != c) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "batch is NULL or foreign");
  int rc = hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (rc) return rc;
  const bool prof = c->profile;
  c->profile = false;  // no event nodes in the graph
  rc = hip_check(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  if (rc) {
    c->profile = prof;
    return rc;
  }
  const int run_rc = mlgi_batch_run(c, b);
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(c->stream, &g);
  c->profile = prof;
  if (run_rc) {
    if (g) (void)hipGraphDestroy(g);
    return run_rc;
  }
  if ((rc = hip_check(c, end, "hipStreamEndCapture"))) return rc;
  auto* og = new mlgi_graph();
  og->graph = g;
  if ((rc = hip_check(c, hipGraphInstantiate(&og->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate"))) {
    (void)hipGraphDestroy(g);
    delete og;
    return rc;
  }
  *graph = og;
  return MLGI_OK;
}

void mlgi_batch_destroy(mlgi_batch* b) {
  if (!b) return;
  if ((b->dev || b->work) && b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
  if (b->dev) (void)hipFree(b->dev);
  if (b->work) (void)hipFree(b->work);
  delete b;
}

}  // extern "C"

// ---------------------------------------------------------------- autotuner

extern "C" int mlgi_autotune(mlgi_context* c, mlgi_frame_desc* frame, int32_t min_vpls_per_pixel,
                             mlgi_tune_result* out) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!frame || !out) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "frame / out is NULL");
  std::memset(out, 0, sizeof *out);
  mlgi_params& p = frame->params;
  const int n = frame->vpls.count;
  const int min_vpp = min_vpls_per_pixel > 0 ? min_vpls_per_pixel : 32;
  int il = 4;
  if (n % 64 == 0 && n / 64 >= min_vpp) il = 8;
  p.interleave = il;
  p.use_interleaving = 1;
  const std::vector<int> key = {frame->gbuffer.width, frame->gbuffer.height, n, p.precision,
                                p.enable_shadowing, il};
  auto it = c->tuned.find(key);
  if (it != c->tuned.end()) {
    p.tile_w = it->second[0];
    p.tile_h = it->second[1];
    out->interleave = il;
    out->tile_w = p.tile_w;
    out->tile_h = p.tile_h;
    out->cached = 1;
    return MLGI_OK;
  }
  static const int kStrip[3] = {8, 16, 32}, kTileH[3] = {8, 4, 8};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = hip_check(c, hipEventCreate(&e0), "hipEventCreate");
  if (!rc) rc = hip_check(c, hipEventCreate(&e1), "hipEventCreate");
  int best = -1;
  float best_ms = 0.0f;
  const bool prof = c->profile;
  c->profile = false;
  for (int k = 0; k < 3 && !rc; ++k) {
    mlgi_frame_desc d = *frame;
    d.params.tile_w = kStrip[k];
    d.params.tile_h = kTileH[k];
    mlgi_batch* b = nullptr;
    if ((rc = mlgi_batch_create(c, &d, 1, &b))) break;
    float ms = 0.0f;
    // one untimed run (fills VPLs and ISM), then the best of three
    rc = mlgi_batch_run(c, b);
    for (int rep = 0; rep < 3 && !rc; ++rep) {
      if ((rc = hip_check(c, hipEventRecord(e0, c->stream), "hipEventRecord"))) break;
      if ((rc = hip_check(c, mlgi::launch_batch_gather(c->stream, b->dev + b->plan.gather[0].begin, b->plan.gather[0].count,
                                                       b->plan.gather[0].total, b->plan.gather[0].a),
                          "batch_gather")))
        break;
      if ((rc = hip_check(c, hipEventRecord(e1, c->stream), "hipEventRecord"))) break;
      if ((rc = hip_check(c, hipEventSynchronize(e1), "hipEventSynchronize"))) break;
      float t = 0.0f;
      if ((rc = hip_check(c, hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime"))) break;
      ms = rep == 0 ? t : std::min(ms, t);
    }
    mlgi_batch_destroy(b);
    if (rc) break;
    out->gather_ms[k] = ms;
    if (best < 0 || ms < best_ms) {
      best = k;
      best_ms = ms;
    }
  }
  c->profile = prof;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc) return rc;
  p.tile_w = kStrip[best];
  p.tile_h = kTileH[best];
  c->tuned[key] = {p.tile_w, p.tile_h};
  out->interleave = il;
  out->tile_w = p.tile_w;
  out->tile_h = p.tile_h;
  return MLGI_OK;
}

// ------------------------------------------------------- device producer

struct mlgi_scene_device {
  mlgi_context* ctx = nullptr;
  mlgi_scene_impl::Node4* nodes4 = nullptr;
  mlgi_scene_impl::TriGeo* geo = nullptr;
  mlgi_scene_impl::Tri* tris = nullptr;
  int* order = nullptr;
  double* cdf = nullptr;  // running triangle areas (point sampling)
  double total_area = 0.0;
  int n_nodes4 = 0, n_tris = 0;
  mlgi_scene_impl::SceneView view() const {
    return mlgi_scene_impl::SceneView{nodes4, geo, tris, order, n_nodes4, n_tris};
  }
};

extern "C" {

void mlgi_scene_device_destroy(mlgi_scene_device* d) {
  if (!d) return;
  if (d->ctx) (void)hipStreamSynchronize(d->ctx->stream);
  (void)hipFree(d->nodes4);
  (void)hipFree(d->geo);
  (void)hipFree(d->tris);
  (void)hipFree(d->order);
  (void)hipFree(d->cdf);
  delete d;
}

int mlgi_scene_upload(mlgi_context* c, const mlgi_scene* s, mlgi_scene_device** out) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!s || !out) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "scene / out is NULL");
  *out = nullptr;
  const mlgi_scene_impl::SceneView v = mlgi_scene_view(s);
  if (v.n_nodes4 <= 0 || v.n_tris <= 0) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "empty scene");
  auto* d = new mlgi_scene_device();
  d->ctx = c;
  d->n_nodes4 = v.n_nodes4;
  d->n_tris = v.n_tris;
  auto copy = [&](auto** dst, const auto* src, size_t n, const char* what) {
    int r = hip_check(c, hipMalloc(dst, sizeof(**dst) * n), what);
    if (!r) r = hip_check(c, hipMemcpy(*dst, src, sizeof(**dst) * n, hipMemcpyHostToDevice), what);
    return r;
  };
  int rc = copy(&d->nodes4, v.nodes4, (size_t)v.n_nodes4, "upload(bvh)");
  if (!rc) rc = copy(&d->geo, v.geo, (size_t)v.n_tris + mlgi_scene_impl::kGeoPad, "upload(triangle geometry)");
  if (!rc) rc = copy(&d->tris, v.tris, (size_t)v.n_tris, "upload(triangles)");
  if (!rc) rc = copy(&d->order, v.order, (size_t)v.n_tris, "upload(order)");
  if (!rc) {
    std::vector<double> cdf(v.n_tris);
    d->total_area = mlgi_scene_area_cdf(s, cdf.data());
    rc = hip_check(c, hipMalloc(&d->cdf, sizeof(double) * v.n_tris), "hipMalloc(cdf)");
    if (!rc)
      rc = hip_check(c, hipMemcpy(d->cdf, cdf.data(), sizeof(double) * v.n_tris, hipMemcpyHostToDevice),
                     "hipMemcpy(cdf)");
  }
  if (rc) {
    mlgi_scene_device_destroy(d);
    return rc;
  }
  *out = d;
  return MLGI_OK;
}

int mlgi_scene_device_sample_points(mlgi_context* c, const mlgi_scene_device* d, int64_t n_points, int32_t n_bins,
                                    int32_t ism_count, uint64_t seed, float* records, uint32_t* bin_offsets) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!d || !bin_offsets || (!records && n_points > 0)) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n_points < 0 || n_points > (1ll << 31) - 1 || n_bins <= 0 || n_bins > (1 << 20) || ism_count <= 0)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "points %lld / bins %d / ism count %d out of range", (long long)n_points,
                n_bins, ism_count);
  const int n = (int)n_points;
  const float r_packed = mlgi_scene_point_radius(d->total_area, n_points, ism_count);
  float* tmp = nullptr;
  int* bin = nullptr;
  uint32_t* cnt = nullptr;
  int rc = hip_check(c, hipMalloc(&tmp, sizeof(float) * 4 * (size_t)std::max(n, 1)), "hipMalloc(points tmp)");
  if (!rc) rc = hip_check(c, hipMalloc(&bin, sizeof(int) * (size_t)std::max(n, 1)), "hipMalloc(points bin)");
  if (!rc) rc = hip_check(c, hipMalloc(&cnt, sizeof(uint32_t) * 2 * (size_t)n_bins), "hipMalloc(points count)");
  if (!rc)
    rc = hip_check(c,
                   mlgi::launch_sample_points(c->stream, d->tris, d->cdf, d->n_tris, d->total_area, seed, r_packed,
                                              n_bins, n, tmp, bin, cnt, cnt + n_bins, records, bin_offsets),
                   "sample_points");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (tmp) (void)hipFree(tmp);
  if (bin) (void)hipFree(bin);
  if (cnt) (void)hipFree(cnt);
  return rc;
}

int mlgi_scene_device_render_surface(mlgi_context* c, const mlgi_scene_device* d, const mlgi_camera* cam, int32_t w,
                                     int32_t h, float* albedo, float* moments) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!d || !cam || (!albedo && !moments)) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "NULL argument");
  if (w <= 0 || h <= 0 || w > 16384 || h > 16384)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "size %dx%d out of range", w, h);
  mlgi_scene_impl::RayCam rc;
  if (!mlgi_scene_ray_cam(cam, &rc)) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "bad camera");
  return hip_check(c, mlgi::launch_raycast(c->stream, d->view(), rc, w, h, 2, albedo, moments, nullptr),
                   "raycast(surface)");
}

int mlgi_scene_device_render_gbuffer(mlgi_context* c, const mlgi_scene_device* d, const mlgi_camera* cam,
                                     int32_t w, int32_t h, float* depth, float* normal, float* vp_inv16) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!d || !cam || !depth || !normal) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "NULL argument");
  if (w <= 0 || h <= 0 || w > 16384 || h > 16384)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "size %dx%d out of range", w, h);
  mlgi_scene_impl::RayCam rc;
  if (!mlgi_scene_ray_cam(cam, &rc)) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "bad camera");
  if (vp_inv16 && mlgi_camera_matrices(cam, nullptr, nullptr, vp_inv16))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "singular camera");
  return hip_check(c, mlgi::launch_raycast(c->stream, d->view(), rc, w, h, 0, depth, normal, nullptr),
                   "raycast(gbuffer)");
}

int mlgi_scene_device_render_rsm(mlgi_context* c, const mlgi_scene_device* d, const mlgi_camera* light,
                                 int32_t w, int32_t h, float* diffuse, float* normal, float* depth,
                                 float* inv_bvp16) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!d || !light || !diffuse || !normal || !depth) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "NULL argument");
  if (w <= 0 || h <= 0 || w > 16384 || h > 16384)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "size %dx%d out of range", w, h);
  mlgi_scene_impl::RayCam rc;
  if (!mlgi_scene_ray_cam(light, &rc)) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "bad light");
  if (inv_bvp16 && !mlgi_scene_inv_biased_vp(light, inv_bvp16))
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "singular light");
  return hip_check(c, mlgi::launch_raycast(c->stream, d->view(), rc, w, h, 1, diffuse, normal, depth),
                   "raycast(rsm)");
}

}  // extern "C"
