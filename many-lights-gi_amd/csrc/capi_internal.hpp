// Internals of the C ABI shared by its translation units (capi.cpp, server.cpp):
// the context, error reporting, stage timing, descriptor validation and the
// batch planner.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/mlgi.h"
#include "kernels.hpp"

struct mlgi_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  std::map<int, int32_t*> shuffle;  // VPL count -> device permutation
  unsigned long long* counter = nullptr;
  // profiling (PerfCounter analogue): event pairs per stage
  bool profile = false;
  struct Span {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Span> spans;
  // Event last recorded at the stream's tail by a Scope, while nothing has
  // been enqueued after it (see Scope); depth of the open scopes.
  hipEvent_t tail = nullptr;
  int scope_depth = 0;
  std::vector<hipEvent_t> pool;
  int frames = 0;
  // autotuner cache: shape key -> (tile_w, tile_h)
  std::map<std::vector<int>, std::vector<int>> tuned;
  // work area of the single-frame ISM-major splat (grown on demand)
  void* splat_ws = nullptr;
  size_t splat_ws_bytes = 0;
  int64_t splat_ws_points = -1;  // points of the last ISM-major splat in splat_ws (-1: none)
  // While mlgi_graph_capture_frame records a frame: that graph's own splat
  // work area, used instead of splat_ws (which a later eager splat may
  // reallocate while the graph still refers to it).
  void* capture_ws = nullptr;
  size_t capture_ws_bytes = 0;
  // The device's static splat work area (mlgi_ism_splat_noalloc), resolved
  // at context creation; shared by every context on the device.
  void* scratch = nullptr;
  // recorded by mlgi_context_wait when another context waits for this one
  hipEvent_t order_event = nullptr;
};

namespace mlgi_capi {

enum Stage { kVpl = 0, kClear, kSplat, kPullPush, kGather, kFrame, kBlur, kStages };

int fail(mlgi_context* c, int code, const char* fmt, ...);
int hip_check(mlgi_context* c, hipError_t e, const char* what);
bool is_pow2(int v);
int log2i(int v);
hipEvent_t take_event(mlgi_context* c);
mlgi::Mat4 to_mat(const float* m);
int get_shuffle(mlgi_context* c, int n, const int32_t** out);
int check_params(mlgi_context* c, const mlgi_params* p);
int interleave_of(const mlgi_params* p);
int check_vpls(mlgi_context* c, const mlgi_vpls* v);
int atlas_geometry(mlgi_context* c, int side, int n_vpl, int* n1d, int* ism_side);
int build_gather_args(mlgi_context* c, const mlgi_gbuffer* gb, const mlgi_vpls* vpls,
                      const mlgi_ism_depth* ism, const mlgi_params* p, mlgi::GatherArgs* a);
void pick_tile(const mlgi_params* p, int* tw, int* th);

// An event at the stream's tail: c->tail when nothing was enqueued since it
// was recorded, else a newly recorded one (nullptr if recording fails).
hipEvent_t tail_event(mlgi_context* c);

// Records a timed span on the context stream when profiling. Adjacent
// boundaries share one event (a stage's end is the next stage's start, an
// outer scope's start its first inner scope's start): every event record
// costs the stream a few microseconds between kernels. c->tail is cleared by
// hip_check, through which every enqueue of the library goes, and when the
// outermost scope ends, so a shared event never spans work enqueued in
// between, nor anything a caller enqueues between two calls. A negative
// stage records nothing (callers that must create no events).
struct Scope {
  mlgi_context* c;
  int stage;
  bool on;
  hipEvent_t a = nullptr;
  Scope(mlgi_context* ctx, int st) : c(ctx), stage(st), on(ctx->profile && st >= 0) {
    if (!on) return;
    ++c->scope_depth;
    a = tail_event(c);
  }
  ~Scope() {
    if (!on) return;
    if (a) {
      hipEvent_t b = tail_event(c);
      if (b) c->spans.push_back({stage, a, b});
    }
    if (--c->scope_depth == 0) c->tail = nullptr;
  }
};

// A batch's launch plan: per-stage item totals and the contiguous runs of
// frames that share one pull-push kernel (ISM side a, levels b) or one gather
// kernel (precision a).
struct BatchPlan {
  int count = 0;
  std::vector<int> row_of;  // caller's frame index -> table row
  uint32_t total_vpl = 0, total_tiles = 0, total_bins = 0;
  int max_ism_side = 0;  // largest ISM side of the batch (tile pass LDS)
  size_t work_bytes = 0;  // splat work areas of all frames (batch_attach_work)
  struct Run {
    int begin, count;
    uint32_t total;
    int a, b;
  };
  std::vector<Run> pullpush, gather;
};

// Validates one frame and fills its table row (first_* unset).
int batch_frame(mlgi_context* c, const mlgi_frame_desc& d, mlgi::BatchFrame* f);
// Validates and orders frames into a device-ready table and its plan.
int batch_plan(mlgi_context* c, const mlgi_frame_desc* frames, int count, std::vector<mlgi::BatchFrame>* table,
               BatchPlan* plan);
// Points the table rows' splat work areas into `base` (plan.work_bytes bytes
// of device memory, one area per frame).
void batch_attach_work(std::vector<mlgi::BatchFrame>* table, char* base);
// Enqueues a planned batch whose table is at `dev` (device memory).
int batch_enqueue(mlgi_context* c, const mlgi::BatchFrame* dev, const BatchPlan& plan);

}  // namespace mlgi_capi
