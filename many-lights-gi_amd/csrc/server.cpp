// GI request server (include/mlgi.h, "serving"): resident scenes, an HBM
// budget arena, continuous batching over the batched kernels.
//
// One host thread drives the server. Each queued-batch slot has a device
// stream of its own (a lane: the server's context for slot 0, a context the
// server creates for each further slot, all of default priority), so batches
// in different slots run beside each other on the device. Each loop
// iteration
//   1. moves requests whose arrival time has passed into the FIFO queue,
//   2. retires finished batches (event query), releasing their slots and
//      arena ranges and recording completion times,
//   3. if a queued-batch slot is free, admits queued requests in FIFO order
//      while in-flight slots and arena space allow, renders each admitted
//      request's RSM on the slot's stream and enqueues them there as one
//      batch.
// Requests arriving while batches run are therefore admitted as soon as
// resources free up, without waiting for the device to drain, and their work
// overlaps the batches already running. An arena range returns to the free
// list only after the host has seen its batch's completion event, so a later
// batch on any stream may reuse it. A request that cannot fit the whole arena
// is rejected.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/mlgi.h"
#include "../../include/mlgi_scene.h"
#include "capi_internal.hpp"
#include "kernels.hpp"

using namespace mlgi_capi;

namespace {

constexpr size_t kAlign = 256;

size_t round_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over the offsets [0, size) of one device allocation.
class Arena {
 public:
  explicit Arena(size_t size) : size_(size) {
    if (size) free_[0] = size;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = round_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, next = it->first + bytes;
      free_.erase(it);
      if (rest) free_[next] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = round_up(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t size() const { return size_; }
  size_t peak() const { return peak_; }
  void reset_peak() { peak_ = used_; }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0, used_ = 0, peak_ = 0;
};

struct GBuf {
  float* depth = nullptr;
  float* normal = nullptr;
  float vp_inv[16];
};

struct PointSet {
  float* records = nullptr;
  uint32_t* offsets = nullptr;
  int64_t count = 0;
};

struct SceneRes {
  const mlgi_scene* host = nullptr;
  mlgi_scene_device* dev = nullptr;
  std::map<int, PointSet> points;                  // VPL count -> binning
  std::map<std::pair<int, int>, GBuf> gbuffers;    // (width, height)
};

struct Tune {
  int interleave, tile_w, tile_h;
};

// Work buffers of one admitted request, carved out of one arena range.
struct Footprint {
  int atlas_side, ism_side, levels;
  size_t vpl, atlas, ism, rsm_diffuse, rsm_normal, rsm_depth, image, work, total;
};

// n_points: the scene's point count (sizes the splat work area, last in the range)
Footprint footprint(int w, int h, int n, int rsm, int n_points) {
  Footprint f;
  f.ism_side = n <= 1024 ? 32 : 64;
  f.atlas_side = mlgi_ism_indices_1d(n) * f.ism_side;
  f.levels = f.ism_side == 32 ? 3 : 4;
  const size_t a2 = (size_t)f.atlas_side * f.atlas_side;
  f.vpl = round_up((size_t)n * MLGI_VPL_RECORD_FLOATS * 4);
  f.atlas = round_up(a2 * 4);
  f.ism = round_up(a2 * 2);
  f.rsm_diffuse = round_up((size_t)rsm * rsm * 12);
  f.rsm_normal = round_up((size_t)rsm * rsm * 12);
  f.rsm_depth = round_up((size_t)rsm * rsm * 4);
  f.image = round_up((size_t)w * h * 12);
  f.work = round_up(mlgi::splat_work_bytes((uint32_t)n_points));
  f.total = f.vpl + f.atlas + f.ism + f.rsm_diffuse + f.rsm_normal + f.rsm_depth + f.image + f.work;
  return f;
}

struct Admitted {
  int req;
  size_t off, bytes;
  float* image;
};

struct Inflight {
  hipEvent_t done;
  int ring;
  std::vector<Admitted> reqs;
};

}  // namespace

struct mlgi_server {
  mlgi_context* ctx = nullptr;
  mlgi_server_config cfg{};
  std::vector<int> vpl_counts;
  std::vector<std::pair<int, int>> resolutions;
  std::vector<SceneRes> scenes;
  std::map<std::tuple<int, int, int>, Tune> tuned;  // (w, h, n)
  char* arena_base = nullptr;
  Arena arena{0};
  // batch tables: ring of max_queued_batches slots of max_batch rows
  mlgi::BatchFrame* table_dev = nullptr;
  mlgi::BatchFrame* table_host = nullptr;  // pinned
  std::vector<hipEvent_t> events;
  std::vector<void*> owned;  // resident device allocations
  // one context (stream) per lane; lanes[0] is ctx, the others are owned
  std::vector<mlgi_context*> lanes;
  mlgi_context* lane_of(int slot) const { return lanes[(size_t)slot % lanes.size()]; }
};

namespace {

int dev_alloc(mlgi_server* s, size_t bytes, void** p) {
  int rc = hip_check(s->ctx, hipMalloc(p, bytes), "hipMalloc(server)");
  if (!rc) s->owned.push_back(*p);
  return rc;
}

int points_of(const mlgi_server* s, const mlgi_request& r) { return s->scenes[r.scene].points.at(r.n_vpl).count; }

mlgi_params request_params(const mlgi_server* s, const Footprint& fp, int n, const Tune& t) {
  mlgi_params p;
  mlgi_params_default(&p);
  p.precision = s->cfg.precision;
  p.atlas_side = fp.atlas_side;
  p.pull_push_levels = fp.levels;
  p.vpl_end_index = n;
  p.z_far = 50.0f;
  p.interleave = t.interleave;
  p.use_interleaving = 1;
  p.tile_w = t.tile_w;
  p.tile_h = t.tile_h;
  return p;
}

// Frame descriptor of request r with its buffers at arena offset `off`.
mlgi_frame_desc request_frame(mlgi_server* s, const mlgi_request& r, size_t off, const Tune& t,
                              const float inv_bvp[16], float** image) {
  const int rs = s->cfg.rsm_size;
  const Footprint fp = footprint(r.width, r.height, r.n_vpl, rs, points_of(s, r));
  char* b = s->arena_base + off;
  float* vpl = (float*)b;
  uint32_t* atlas = (uint32_t*)(b + fp.vpl);
  uint16_t* ism = (uint16_t*)(b + fp.vpl + fp.atlas);
  float* dif = (float*)(b + fp.vpl + fp.atlas + fp.ism);
  float* nrm = (float*)((char*)dif + fp.rsm_diffuse);
  float* dep = (float*)((char*)nrm + fp.rsm_normal);
  float* img = (float*)((char*)dep + fp.rsm_depth);
  SceneRes& sc = s->scenes[r.scene];
  const PointSet& ps = sc.points.at(r.n_vpl);
  const GBuf& gb = sc.gbuffers.at({r.width, r.height});
  mlgi_frame_desc d;
  std::memset(&d, 0, sizeof d);
  d.rsm = mlgi_rsm{dif, nrm, dep, rs, rs, {}};
  std::memcpy(d.rsm.inv_biased_view_projection, inv_bvp, 64);
  d.points = mlgi_points{ps.records, ps.offsets, (int32_t)ps.count, r.n_vpl};
  d.gbuffer = mlgi_gbuffer{gb.depth, gb.normal, r.width, r.height, {}};
  std::memcpy(d.gbuffer.view_projection_inverse, gb.vp_inv, 64);
  d.params = request_params(s, fp, r.n_vpl, t);
  d.vpls = mlgi_vpls{vpl, r.n_vpl};
  d.atlas = mlgi_atlas{atlas, fp.atlas_side, fp.ism_side};
  d.ism = mlgi_ism_depth{ism, fp.atlas_side, fp.ism_side};
  d.image = mlgi_image{img, r.width, r.height};
  if (image) *image = img;
  return d;
}

// Renders the request's RSM on lane L's stream (asynchronous) and returns its
// inverse biased light matrix.
int render_request_rsm(mlgi_server* s, mlgi_context* L, const mlgi_request& r, size_t off, float inv_bvp[16]) {
  const int rs = s->cfg.rsm_size;
  const Footprint fp = footprint(r.width, r.height, r.n_vpl, rs, points_of(s, r));
  char* b = s->arena_base + off;
  float* dif = (float*)(b + fp.vpl + fp.atlas + fp.ism);
  float* nrm = (float*)((char*)dif + fp.rsm_diffuse);
  float* dep = (float*)((char*)nrm + fp.rsm_normal);
  mlgi_camera light;
  mlgi_scene_light_pose(r.light_seed, &light);
  return mlgi_scene_device_render_rsm(L, s->scenes[r.scene].dev, &light, rs, rs, dif, nrm, dep, inv_bvp);
}

// Copies a lane's error into the server context's (no-op for lane 0).
int lane_fail(mlgi_server* s, mlgi_context* L, int rc) {
  if (rc && L != s->ctx) s->ctx->error = L->error;
  return rc;
}

bool configured(const std::vector<int>& v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); }

}  // namespace

extern "C" {

void mlgi_server_destroy(mlgi_server* s) {
  if (!s) return;
  for (mlgi_context* L : s->lanes) (void)hipStreamSynchronize(L->stream);
  for (size_t k = 1; k < s->lanes.size(); ++k) mlgi_context_destroy(s->lanes[k]);
  if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
  for (auto e : s->events) (void)hipEventDestroy(e);
  for (auto& sc : s->scenes) mlgi_scene_device_destroy(sc.dev);
  for (void* p : s->owned) (void)hipFree(p);
  if (s->table_host) (void)hipHostFree(s->table_host);
  delete s;
}

int mlgi_server_create(mlgi_context* c, const mlgi_server_config* cfg, const mlgi_server_scene* scenes,
                       int32_t n_scenes, mlgi_server** out) {
  if (!c) return MLGI_ERR_INVALID_ARGUMENT;
  if (!cfg || !scenes || !out || n_scenes <= 0) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (cfg->max_inflight <= 0 || cfg->max_batch <= 0 || cfg->max_queued_batches <= 0 || cfg->rsm_size <= 0 ||
      cfg->rsm_size > 4096 || cfg->n_vpl_counts <= 0 || !cfg->vpl_counts || cfg->n_resolutions <= 0 ||
      !cfg->resolutions || cfg->hbm_budget == 0 || cfg->streams < 0 || cfg->max_queued_batches > 64)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "bad server configuration");
  if (cfg->precision != MLGI_PRECISION_FP32 && cfg->precision != MLGI_PRECISION_MIXED)
    return fail(c, MLGI_ERR_INVALID_ARGUMENT, "unknown precision %d", cfg->precision);
  auto* s = new mlgi_server();
  s->ctx = c;
  s->cfg = *cfg;
  s->vpl_counts.assign(cfg->vpl_counts, cfg->vpl_counts + cfg->n_vpl_counts);
  for (int i = 0; i < cfg->n_resolutions; ++i)
    s->resolutions.push_back({cfg->resolutions[2 * i], cfg->resolutions[2 * i + 1]});
  s->cfg.vpl_counts = nullptr;
  s->cfg.resolutions = nullptr;
  int rc = MLGI_OK;
  // lanes: the server's context and one more context (stream) per further slot
  const int n_lanes = std::min(cfg->streams > 0 ? cfg->streams : cfg->max_queued_batches, cfg->max_queued_batches);
  s->lanes.push_back(c);
  for (int k = 1; k < n_lanes && !rc; ++k) {
    mlgi_context* L = nullptr;
    const int r = mlgi_context_create(c->device, &L);
    if (r) rc = fail(c, r, "server lane context %d could not be created", k);
    else s->lanes.push_back(L);
  }
  for (int n : s->vpl_counts) {
    if (rc) break;
    if (n <= 0 || n % 16 != 0) rc = fail(c, MLGI_ERR_INVALID_ARGUMENT, "VPL count %d must be a multiple of 16", n);
    for (mlgi_context* L : s->lanes) {  // each lane's shuffle table, before any request
      const int32_t* sh = nullptr;
      if (!rc) rc = lane_fail(s, L, get_shuffle(L, n, &sh));
    }
  }
  for (auto& wh : s->resolutions)
    if (!rc && (wh.first <= 0 || wh.second <= 0 || wh.first > 16384 || wh.second > 16384))
      rc = fail(c, MLGI_ERR_INVALID_ARGUMENT, "resolution %dx%d out of range", wh.first, wh.second);
  // resident scenes: BVH, point binnings per VPL count, G-buffers per resolution
  for (int i = 0; i < n_scenes && !rc; ++i) {
    SceneRes sc;
    sc.host = scenes[i].scene;
    if (!sc.host || scenes[i].n_points < 0) {
      rc = fail(c, MLGI_ERR_INVALID_ARGUMENT, "scene %d is NULL", i);
      break;
    }
    s->scenes.push_back(sc);
    SceneRes& r = s->scenes.back();
    if ((rc = mlgi_scene_upload(c, r.host, &r.dev))) break;
    for (int n : s->vpl_counts) {
      PointSet ps;
      ps.count = scenes[i].n_points;
      std::vector<float> rec((size_t)ps.count * 4 + 4);
      std::vector<uint32_t> off((size_t)n + 1);
      if (mlgi_scene_sample_points(r.host, ps.count, n, n, scenes[i].point_seed, 0, rec.data(), off.data())) {
        rc = fail(c, MLGI_ERR_INVALID_ARGUMENT, "point sampling failed");
        break;
      }
      if ((rc = dev_alloc(s, rec.size() * 4, (void**)&ps.records))) break;
      if ((rc = dev_alloc(s, off.size() * 4, (void**)&ps.offsets))) break;
      if ((rc = hip_check(c, hipMemcpy(ps.records, rec.data(), rec.size() * 4, hipMemcpyHostToDevice), "hipMemcpy")))
        break;
      if ((rc = hip_check(c, hipMemcpy(ps.offsets, off.data(), off.size() * 4, hipMemcpyHostToDevice), "hipMemcpy")))
        break;
      r.points[n] = ps;
    }
    for (auto& wh : s->resolutions) {
      if (rc) break;
      GBuf g;
      const size_t px = (size_t)wh.first * wh.second;
      if ((rc = dev_alloc(s, px * 4, (void**)&g.depth)) || (rc = dev_alloc(s, px * 12, (void**)&g.normal))) break;
      mlgi_camera cam;
      mlgi_scene_default_camera((float)wh.first / (float)wh.second, &cam);
      if ((rc = mlgi_scene_device_render_gbuffer(c, r.dev, &cam, wh.first, wh.second, g.depth, g.normal, g.vp_inv)))
        break;
      r.gbuffers[wh] = g;
    }
  }
  // arena, batch table ring, events
  if (!rc) rc = dev_alloc(s, cfg->hbm_budget, (void**)&s->arena_base);
  if (!rc) s->arena = Arena(cfg->hbm_budget);
  const size_t ring_rows = (size_t)cfg->max_batch * cfg->max_queued_batches;
  if (!rc) rc = dev_alloc(s, sizeof(mlgi::BatchFrame) * ring_rows, (void**)&s->table_dev);
  if (!rc)
    rc = hip_check(c, hipHostMalloc((void**)&s->table_host, sizeof(mlgi::BatchFrame) * ring_rows, hipHostMallocDefault),
                   "hipHostMalloc");
  for (int i = 0; i < cfg->max_queued_batches && !rc; ++i) {
    hipEvent_t e = nullptr;
    rc = hip_check(c, hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    if (!rc) s->events.push_back(e);
  }
  // tune every shape on scene 0 with a fixed light (the autotuner caches per shape)
  for (auto& wh : s->resolutions) {
    for (int n : s->vpl_counts) {
      if (rc) break;
      mlgi_request r{0, wh.first, wh.second, n, 0, 0.0, nullptr};
      const Footprint fp = footprint(wh.first, wh.second, n, cfg->rsm_size, points_of(s, r));
      if (fp.total > cfg->hbm_budget) continue;  // never admissible; requests of it are rejected
      size_t off = 0;
      s->arena.alloc(fp.total, &off);
      float inv[16];
      if ((rc = render_request_rsm(s, c, r, off, inv))) break;
      Tune t{4, 8, 8};
      mlgi_frame_desc d = request_frame(s, r, off, t, inv, nullptr);
      mlgi_tune_result tr;
      if ((rc = mlgi_autotune(c, &d, cfg->min_vpls_per_pixel, &tr))) break;
      s->tuned[{wh.first, wh.second, n}] = Tune{tr.interleave, tr.tile_w, tr.tile_h};
      rc = hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
      s->arena.release(off, fp.total);
    }
  }
  if (rc) {
    std::string m = c->error;
    mlgi_server_destroy(s);
    return fail(c, rc, "server create: %s", m.c_str());
  }
  s->arena.reset_peak();
  *out = s;
  return MLGI_OK;
}

int mlgi_server_run(mlgi_server* s, const mlgi_request* reqs, int32_t n, mlgi_server_stats* stats,
                    double* latencies_ms) {
  if (!s) return MLGI_ERR_INVALID_ARGUMENT;
  mlgi_context* c = s->ctx;
  if ((!reqs && n) || n < 0 || !stats) return fail(c, MLGI_ERR_INVALID_ARGUMENT, "NULL argument");
  for (int i = 0; i < n; ++i) {
    const mlgi_request& r = reqs[i];
    if (r.scene < 0 || r.scene >= (int)s->scenes.size() || !configured(s->vpl_counts, r.n_vpl) ||
        std::find(s->resolutions.begin(), s->resolutions.end(), std::make_pair(r.width, r.height)) ==
            s->resolutions.end())
      return fail(c, MLGI_ERR_INVALID_ARGUMENT, "request %d: scene / resolution / VPL count not configured", i);
    if (i > 0 && r.arrival_s < reqs[i - 1].arrival_s)
      return fail(c, MLGI_ERR_INVALID_ARGUMENT, "requests must be sorted by arrival");
  }
  std::memset(stats, 0, sizeof *stats);
  std::vector<double> lat(n, -1.0);
  std::deque<int> queue;
  std::deque<Inflight> inflight;
  std::vector<int> ring_free;
  for (int i = (int)s->events.size() - 1; i >= 0; --i) ring_free.push_back(i);
  int next = 0, done = 0, inflight_reqs = 0;
  int64_t batched = 0;
  const auto t0 = std::chrono::steady_clock::now();
  auto now_s = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  double last_done = 0.0;
  int rc = MLGI_OK;
  while (done < n && !rc) {
    bool progress = false;
    const double t = now_s();
    // 1. arrivals
    while (next < n && reqs[next].arrival_s <= t) queue.push_back(next++);
    // 2. completions
    while (!inflight.empty()) {
      const hipError_t q = hipEventQuery(inflight.front().done);
      if (q == hipErrorNotReady) break;
      if ((rc = hip_check(c, q, "hipEventQuery"))) break;
      const double tc = now_s();
      for (const Admitted& a : inflight.front().reqs) {
        lat[a.req] = (tc - reqs[a.req].arrival_s) * 1e3;
        s->arena.release(a.off, a.bytes);
        --inflight_reqs;
        ++done;
      }
      last_done = tc;
      ring_free.push_back(inflight.front().ring);
      inflight.pop_front();
      progress = true;
    }
    if (rc) break;
    // 3. admission into one new batch, on the stream of the next free slot
    if (!queue.empty() && !ring_free.empty()) {
      const int slot = ring_free.back();
      mlgi_context* L = s->lane_of(slot);
      std::vector<Admitted> adm;
      std::vector<mlgi_frame_desc> descs;
      while (!queue.empty() && (int)adm.size() < s->cfg.max_batch && inflight_reqs < s->cfg.max_inflight) {
        const int i = queue.front();
        const mlgi_request& r = reqs[i];
        const Footprint fp = footprint(r.width, r.height, r.n_vpl, s->cfg.rsm_size, points_of(s, r));
        if (fp.total > s->arena.size()) {  // can never fit: reject
          queue.pop_front();
          stats->rejected++;
          ++done;
          progress = true;
          continue;
        }
        size_t off = 0;
        if (!s->arena.alloc(fp.total, &off)) break;  // wait for releases (FIFO, no overtaking)
        queue.pop_front();
        float inv[16];
        if ((rc = lane_fail(s, L, render_request_rsm(s, L, r, off, inv)))) break;
        float* img = nullptr;
        descs.push_back(request_frame(s, r, off, s->tuned.at({r.width, r.height, r.n_vpl}), inv, &img));
        adm.push_back({i, off, fp.total, img});
        ++inflight_reqs;
      }
      if (!rc && !adm.empty()) {
        ring_free.pop_back();
        std::vector<mlgi::BatchFrame> table;
        BatchPlan plan;
        if (!(rc = lane_fail(s, L, batch_plan(L, descs.data(), (int)descs.size(), &table, &plan)))) {
          for (size_t k = 0; k < adm.size(); ++k) {  // splat work areas: the end of each request's range
            const mlgi_request& r = reqs[adm[k].req];
            const Footprint fp = footprint(r.width, r.height, r.n_vpl, s->cfg.rsm_size, points_of(s, r));
            mlgi::BatchFrame& row = table[plan.row_of[k]];
            const mlgi::SplatWork w = mlgi::splat_work(s->arena_base + adm[k].off + fp.total - fp.work, row.n_points);
            row.mask = w.mask;
            row.spill_state = w.state;
            row.spill = w.spill;
          }
          mlgi::BatchFrame* host = s->table_host + (size_t)slot * s->cfg.max_batch;
          mlgi::BatchFrame* dev = s->table_dev + (size_t)slot * s->cfg.max_batch;
          std::memcpy(host, table.data(), sizeof(mlgi::BatchFrame) * table.size());
          rc = hip_check(c, hipMemcpyAsync(dev, host, sizeof(mlgi::BatchFrame) * table.size(), hipMemcpyHostToDevice,
                                           L->stream),
                         "hipMemcpyAsync(batch table)");
          if (!rc) rc = lane_fail(s, L, batch_enqueue(L, dev, plan));
          for (const Admitted& a : adm) {
            if (rc || !reqs[a.req].host_out) continue;
            const size_t bytes = (size_t)reqs[a.req].width * reqs[a.req].height * 12;
            rc = hip_check(c, hipMemcpyAsync(reqs[a.req].host_out, a.image, bytes, hipMemcpyDeviceToHost, L->stream),
                           "hipMemcpyAsync(result)");
          }
          if (!rc) rc = hip_check(c, hipEventRecord(s->events[slot], L->stream), "hipEventRecord");
        }
        if (!rc) {
          if (!inflight.empty()) stats->overlapped_batches++;
          inflight.push_back({s->events[slot], slot, std::move(adm)});
          stats->batches++;
          batched += (int64_t)inflight.back().reqs.size();
          stats->max_inflight_seen = std::max(stats->max_inflight_seen, inflight_reqs);
          progress = true;
        }
      }
    }
    if (!progress) std::this_thread::yield();
  }
  for (mlgi_context* L : s->lanes) (void)hipStreamSynchronize(L->stream);
  if (rc) return rc;
  std::vector<double> ok;
  for (int i = 0; i < n; ++i)
    if (lat[i] >= 0.0) ok.push_back(lat[i]);
  std::sort(ok.begin(), ok.end());
  auto pct = [&](double q) {
    if (ok.empty()) return 0.0;
    const size_t k = (size_t)std::min<double>((double)ok.size() - 1, std::ceil(q * ok.size()) - 1);
    return ok[k];
  };
  stats->completed = (int64_t)ok.size();
  stats->wall_s = last_done;
  stats->req_per_s = last_done > 0 ? (double)ok.size() / last_done : 0.0;
  stats->latency_p50_ms = pct(0.50);
  stats->latency_p99_ms = pct(0.99);
  double sum = 0.0;
  for (double v : ok) sum += v;
  stats->latency_mean_ms = ok.empty() ? 0.0 : sum / ok.size();
  stats->latency_max_ms = ok.empty() ? 0.0 : ok.back();
  stats->mean_batch = stats->batches ? (double)batched / stats->batches : 0.0;
  stats->peak_bytes = s->arena.peak();
  stats->streams = (int32_t)s->lanes.size();
  if (latencies_ms) std::copy(lat.begin(), lat.end(), latencies_ms);
  return MLGI_OK;
}

}  // extern "C"
