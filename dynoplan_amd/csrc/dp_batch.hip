// dp_batch.hip -- many independent planning instances in one launch.
//
// A batch holds n expansion contexts (any mix of systems on one device).
// Each instance's pipeline up to its conflict lists runs on its own stream,
// forked from a master stream.  The ordered walks, the latency-bound stage of
// each instance (a chain of dependent nodes that keeps few waves busy), then
// run as one work queue per system class on the master stream: waves claim
// (instance, node) items interleaved over the instances, so the chains of
// all instances advance together (dp_walk_queue, k_grid.hip).  In graph mode
// the whole batch is captured once into one hipGraph and replayed by a single
// hipGraphLaunch while nothing the graph bakes in has changed (tracked by a
// global state epoch that every reallocation and every change of nodes
// count, primitives, world, visited set or surrogate bumps).  In stream mode
// the same work is launched directly.  A one-thread kernel at the start of
// the batch stamps the GPU's constant-rate clock, and the wave that finishes
// an instance's last node stamps it again, which gives each instance's
// completion latency without host round trips.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <vector>

#include "dp_internal.h"

static std::atomic<uint64_t> g_epoch{1};
void dp_bump_epoch() { g_epoch.fetch_add(1, std::memory_order_relaxed); }
uint64_t dp_epoch() { return g_epoch.load(std::memory_order_relaxed); }

namespace {
__global__ void k_stamp(unsigned long long* slot) { *slot = (unsigned long long)wall_clock64(); }
}  // namespace

struct dp_batch {
  std::vector<dp_ctx*> ctx;
  int device = 0;
  int32_t mode = DP_BATCH_GRAPH;
  hipStream_t stream = nullptr;  // master: start stamp, fork, join
  hipStream_t last = nullptr;    // stream of the last launch
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t epoch = 0;
  std::vector<dp_expand_params> params;
  unsigned long long* d_stamp = nullptr;  // [n] instance completions, [n] batch start
  dp_walk_queue* walk = nullptr;
  int wall_khz = 0;
  bool stamped = false;
  int64_t captures = 0, launches = 0;
};

#define B_HIP(expr)                                                          \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      dp_set_error(std::string(#expr) + ": " + hipGetErrorString(e_));       \
      return DP_ERR_HIP;                                                     \
    }                                                                        \
  } while (0)

static void release_graph(dp_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static int check_params(const dp_ctx* c, const dp_expand_params* p) {
  if (!(p->alpha > 0.0f && p->alpha < 1.0f)) {  // dbastar.cpp:588-590
    dp_set_error("Alpha needs to be between 0 and 1!");
    return DP_ERR_ARG;
  }
  if (!(p->delta >= 0.0f) || p->num_check_goal < 0 || p->num_check_goal > 8) {
    dp_set_error("invalid delta or num_check_goal");
    return DP_ERR_ARG;
  }
  if (c->P <= 0 || !c->have_world || c->N <= 0) {
    dp_set_error("every context of a batch needs primitives, a world and nodes");
    return DP_ERR_STATE;
  }
  return DP_OK;
}

// Host-side preparation of every instance (may allocate and synchronise):
// check_goal indices, pair buffers sized from the exact applicable count.
static int prepare(dp_batch* b, const dp_expand_params* params) {
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    dp_ctx* c = b->ctx[i];
    int rc = check_params(c, &params[i]);
    if (rc) return rc;
    if ((rc = dp_prepare_goal_index(c, &params[i], c->stream))) return rc;
    if ((rc = dp_size_pairs(c, &params[i], c->stream))) return rc;
  }
  return DP_OK;
}

// Start stamp, fork, one pipeline per instance on its own stream up to its
// conflict lists, join back into the master stream, then the walk queue.
static int launch_branches(dp_batch* b, const dp_expand_params* params, bool capturing) {
  const int32_t n = (int32_t)b->ctx.size();
  k_stamp<<<1, 1, 0, b->stream>>>(b->d_stamp + n);
  B_HIP(hipGetLastError());
  B_HIP(hipEventRecord(b->fork, b->stream));
  for (int32_t i = 0; i < n; ++i) {
    dp_ctx* c = b->ctx[i];
    B_HIP(hipStreamWaitEvent(c->stream, b->fork, 0));
    c->no_events = capturing;
    c->defer_walk = true;
    int rc = dp_launch_expand(c, &params[i], c->stream);
    c->no_events = false;
    c->defer_walk = false;
    if (rc) return rc;
    c->last_pairs = -1;
    B_HIP(hipEventRecord(b->join[i], c->stream));
    B_HIP(hipStreamWaitEvent(b->stream, b->join[i], 0));
  }
  // the walk arguments are known once every pipeline is launched; a capture
  // reuses those of the eager pass before it
  int rc = capturing ? DP_OK : dp_walk_queue_prepare(b->walk, b->ctx.data(), b->last);
  if (!rc) rc = dp_walk_queue_launch(b->walk, b->d_stamp, b->stream);
  return rc;
}

static int capture(dp_batch* b, const dp_expand_params* params) {
  release_graph(b);
  // One eager pass first: it allocates whatever the pipelines need and
  // rebuilds stale visited grids, so the capture below records kernels only.
  int rc = prepare(b, params);
  if (rc) return rc;
  rc = launch_branches(b, params, false);
  if (rc) return rc;
  B_HIP(hipStreamSynchronize(b->stream));
  const uint64_t epoch = dp_epoch();
  B_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  rc = launch_branches(b, params, true);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (!rc) dp_set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    return rc ? rc : DP_ERR_HIP;
  }
  if (dp_epoch() != epoch) {
    (void)hipGraphDestroy(g);
    dp_set_error("device state changed while the batch graph was captured");
    return DP_ERR_STATE;
  }
  b->graph = g;
  B_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->epoch = epoch;
  b->params.assign(params, params + b->ctx.size());
  ++b->captures;
  return DP_OK;
}

extern "C" {

dp_batch* dp_batch_create(dp_ctx* const* ctxs, int32_t n, int32_t mode) {
  if (!ctxs || n <= 0 || (mode != DP_BATCH_GRAPH && mode != DP_BATCH_STREAMS)) {
    dp_set_error("invalid argument");
    return nullptr;
  }
  for (int32_t i = 0; i < n; ++i) {
    if (!ctxs[i] || ctxs[i]->device != ctxs[0]->device) {
      dp_set_error("a batch needs non-null contexts on one device");
      return nullptr;
    }
    for (int32_t k = 0; k < i; ++k)
      if (ctxs[k] == ctxs[i]) {
        dp_set_error("a context appears twice in the batch");
        return nullptr;
      }
  }
  dp_batch* b = new dp_batch;
  b->ctx.assign(ctxs, ctxs + n);
  b->device = ctxs[0]->device;
  b->mode = mode;
  b->join.assign(n, nullptr);
  bool ok = hipSetDevice(b->device) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
  for (int32_t i = 0; i < n && ok; ++i) ok = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) == hipSuccess;
  ok = ok && hipMalloc(&b->d_stamp, sizeof(unsigned long long) * (n + 1)) == hipSuccess;
  ok = ok && (b->walk = dp_walk_queue_create(n)) != nullptr;
  ok = ok && hipDeviceGetAttribute(&b->wall_khz, hipDeviceAttributeWallClockRate, b->device) == hipSuccess;
  if (!ok || b->wall_khz <= 0) {
    dp_set_error("batch creation failed (stream, events, stamps or clock rate)");
    dp_batch_destroy(b);
    return nullptr;
  }
  return b;
}

void dp_batch_destroy(dp_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  release_graph(b);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (hipEvent_t e : b->join)
    if (e) (void)hipEventDestroy(e);
  if (b->d_stamp) (void)hipFree(b->d_stamp);
  dp_walk_queue_destroy(b->walk);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int dp_batch_expand(dp_batch* b, const dp_expand_params* params, void* stream) {
  if (!b || !params) {
    dp_set_error("null argument");
    return DP_ERR_ARG;
  }
  B_HIP(hipSetDevice(b->device));
  const size_t n = b->ctx.size();
  int rc;
  b->stamped = false;
  if (b->mode == DP_BATCH_STREAMS) {
    if ((rc = prepare(b, params))) return rc;
    if (stream) B_HIP(hipStreamSynchronize((hipStream_t)stream));
    if ((rc = launch_branches(b, params, false))) return rc;
    b->last = b->stream;
  } else {
    const bool same = b->exec && b->epoch == dp_epoch() && b->params.size() == n &&
                      std::memcmp(b->params.data(), params, sizeof(dp_expand_params) * n) == 0;
    if (!same && (rc = capture(b, params))) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : b->stream;
    B_HIP(hipGraphLaunch(b->exec, s));
    b->last = s;
  }
  ++b->launches;
  b->stamped = true;
  return DP_OK;
}

int dp_batch_latency(dp_batch* b, float* ms) {
  if (!b || !ms) {
    dp_set_error("null argument");
    return DP_ERR_ARG;
  }
  if (!b->stamped) {
    dp_set_error("no batch has been launched");
    return DP_ERR_STATE;
  }
  const size_t n = b->ctx.size();
  B_HIP(hipSetDevice(b->device));
  B_HIP(hipStreamSynchronize(b->last));
  std::vector<unsigned long long> t(n + 1);
  B_HIP(hipMemcpy(t.data(), b->d_stamp, sizeof(unsigned long long) * (n + 1), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) ms[i] = (float)((double)(long long)(t[i] - t[n]) / (double)b->wall_khz);
  return DP_OK;
}

int dp_batch_stats(const dp_batch* b, int64_t* captures, int64_t* launches) {
  if (!b) {
    dp_set_error("null argument");
    return DP_ERR_ARG;
  }
  if (captures) *captures = b->captures;
  if (launches) *launches = b->launches;
  return DP_OK;
}

}  // extern "C"
