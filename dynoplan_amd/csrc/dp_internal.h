// dp_internal.h -- context layout shared by the C-ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "dp_device.h"

// Largest obstacle count staged in LDS by one workgroup.
#define DP_MAX_OBS 1024
// obstacles listed per frontier node for the swept-box cull of its pairs
#define DP_NODE_OBS 128
// Per-pair resolution value while undecided (ordered duplicate resolution).
#define DP_R_UNDECIDED 0x7f
// Pair flag set by the pre-resolution: every earlier contender is decided
// and not NOVEL, so the pair is NOVEL unless its node's cap falls before it.
#define DP_F_NUC 0x10
// Pair flag (quadrotor): a FREE pair past its node's FP64 window; its FP64
// rollout, final state, goal index, g and h are computed only if the node's
// cap does not fall before it (the batch is then resolved again).
#define DP_F_DEFER 0x20
// Pair flag (limit_branching_factor == 0): a FREE pair after its node's
// first FREE pair, CAPPED before the walk.
#define DP_F_PRECAP 0x40
// Pair flag (planar systems, limit_branching_factor > 0): a FREE pair past
// its node's list window.  It stays a contender (later pairs test against
// it) but gets no stored list of earlier conflicts; the walk scans the
// contender grid for it if its node's cap does not fall before it.  (The
// quadrotor's check stage uses this bit for its own state before it writes
// the final flags, and never sets it for the walk.)
#define DP_F_NOLIST 0x08
// FP64 window of a quadrotor node, in FREE pairs: enough for the
// cap of limit NOVEL pairs with a wide margin for duplicates.
#define DP_FP64_WINDOW(limit) (3 * (limit) + 32)
// Pad added to swept boxes of the broad phase (m): bounds the rounding of
// "canonical sweep + offset" versus per-state "state + offset" by 1e6x.
#define DP_SWEEP_PAD 1e-9
// Bound on back-off iterations of one wave in the ticketed resolution;
// reaching it leaves the node to the host-driven rounds instead of waiting.
#define DP_SPIN_LIMIT (1 << 20)
// Persistent resolution: polls (s_sleep 1 each) one lane may spend waiting
// for an undecided contender of an earlier node, and the retries of a node
// whose waits all ran out (bounds the kernel's run time if a wait never ends).
#define DP_WAIT_POLLS (1 << 14)
// Resolution trace, per node: start and end clock of its wave, cycles
// waiting, evaluating lists and walking chunks, chunks, waits, polls.
#define DP_TRACE_FIELDS 8
// HIP events of one expansion (stage and kernel timing)
#define DP_EVENTS 8
// LDS per workgroup on gfx950
#define DP_LDS_MAX 163840
#define DP_SPIN_LIMIT_WAITING (1 << 10)
// Earlier-conflict entries stored inline per pair (longer lists continue in
// the overflow pool).
#define DP_CONF_SLOTS 64
// longest stored earlier-conflict list (inline slots + overflow pool)
#define DP_CONF_MAX 4096

void dp_set_error(const std::string& msg);

// Device-side counters, one slot per dp_counts field (int64 each) plus flags.
enum {
  DPC_PAIRS = 0, DPC_BOUNDS, DPC_COLLIDE, DPC_FREE, DPC_SURR, DPC_NOVEL, DPC_DUPV, DPC_DUPB,
  DPC_CAPPED, DPC_GOAL, DPC_CHECKED, DPC_RECHECK, DPC_ROUNDS, DPC_ERRFLAG, DPC_OVERFLOW,
  DPC_NODES_LEFT, DPC_CONF, DPC_TICKET, DPC_CELLS, DPC_WINDOW_OVF, DPC_DEFERRED, DPC_Q64, DPC_CONSTRAINT,
  DPC_COUNT
};
// Sticky counters (never reset by a batch): resolutions finished by
// host-launched rounds after a wave gave up, FP64 window overflows.
enum { DPS_GIVEUP = 0, DPS_WINDOW = 1, DPS_COUNT = 4 };

// Uniform grid used by the duplicate tests.  Axes 0,1 are the position
// (x, y); axis 2 is z for 3-D systems and the heading theta (state[2],
// periodic) for planar ones.  Cells are at least r/w0 wide on the position
// axes and r/w1 on the heading axis, so every state within metric distance
// r of a query lies in the 3x3x3 block of cells around the query's cell.
struct dp_grid_geom {
  double lo[3];
  double inv_cell[3];
  int32_t dims[3];
  int32_t periodic2;
  int64_t ncells;
};

__host__ __device__ inline int32_t dp_cell_axis(double x, double lo, double inv_cell, int32_t dim) {
  double f = (x - lo) * inv_cell;
  int32_t c = f < 0.0 ? 0 : (f >= (double)(dim - 1) ? dim - 1 : (int32_t)f);
  return c;
}

__host__ __device__ inline int32_t dp_cell_theta(double th, double inv_cell, int32_t dim) {
  double t = th;
  if (!(t >= -DP_PI && t < DP_PI)) t = t - DP_TWO_PI * floor((t + DP_PI) / DP_TWO_PI);
  return dp_cell_axis(t, -DP_PI, inv_cell, dim);
}

// Base cell coordinates of state x (components 0, 1, 2).
__host__ __device__ inline void dp_cell_coords(const dp_grid_geom& G, const double* x, int32_t* c) {
  c[0] = dp_cell_axis(x[0], G.lo[0], G.inv_cell[0], G.dims[0]);
  c[1] = dp_cell_axis(x[1], G.lo[1], G.inv_cell[1], G.dims[1]);
  c[2] = G.periodic2 ? dp_cell_theta(x[2], G.inv_cell[2], G.dims[2])
                     : dp_cell_axis(x[2], G.lo[2], G.inv_cell[2], G.dims[2]);
}

__host__ __device__ inline int64_t dp_cell_linear(const dp_grid_geom& G, const int32_t* c) {
  return ((int64_t)c[2] * G.dims[1] + c[1]) * G.dims[0] + c[0];
}

// Neighbour cq (0..26) of base cell c, or -1 when outside the grid (or a
// repeat of the base cell on a single-bin heading axis).
__host__ __device__ inline int64_t dp_cell_neighbor(const dp_grid_geom& G, const int32_t* c, int cq) {
  const int32_t d0 = cq % 3 - 1, d1 = (cq / 3) % 3 - 1, d2 = cq / 9 - 1;
  const int32_t n0 = c[0] + d0, n1 = c[1] + d1;
  if (n0 < 0 || n0 >= G.dims[0] || n1 < 0 || n1 >= G.dims[1]) return -1;
  int32_t n2 = c[2] + d2;
  if (G.dims[2] == 1) {
    if (d2 != 0) return -1;
    n2 = 0;
  } else if (G.periodic2) {
    n2 = n2 < 0 ? n2 + G.dims[2] : (n2 >= G.dims[2] ? n2 - G.dims[2] : n2);
  } else if (n2 < 0 || n2 >= G.dims[2]) {
    return -1;
  }
  return ((int64_t)n2 * G.dims[1] + n1) * G.dims[0] + n0;
}

// Learned surrogate state (k_mlp.hip).
struct dp_mlp {
  int32_t P = 0;
  int64_t step = 0;
  bool ready = false;        // trained and calibrated
  bool cache_valid = false;  // per-primitive first-layer terms up to date
  float threshold = -INFINITY;
  float* w = nullptr;  int64_t w_cap = 0;    // FP32 master parameters
  float* g = nullptr;  int64_t g_cap = 0;    // gradients
  float* m1 = nullptr; int64_t m1_cap = 0;   // Adam moments
  float* m2 = nullptr; int64_t m2_cap = 0;
  uint16_t* wb = nullptr; int64_t wb_cap = 0;  // BF16 parameters
  float* emb = nullptr; int64_t emb_cap = 0;   // [P][64] x {value, grad, m, v}
  uint16_t* embb = nullptr; int64_t embb_cap = 0;
  // activations / scratch of training
  float *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *dh = nullptr, *zb = nullptr;
  int64_t h1_cap = 0, h2_cap = 0, h3_cap = 0, dh_cap = 0, zb_cap = 0;
  uint16_t *h1b = nullptr, *h2b = nullptr, *h3b = nullptr, *dhb = nullptr, *xb = nullptr;
  int64_t h1b_cap = 0, h2b_cap = 0, h3b_cap = 0, dhb_cap = 0, xb_cap = 0;
  // inference
  uint16_t* patch = nullptr; int64_t patch_cap = 0;  // [N][1024]
  float* An = nullptr; int64_t An_cap = 0;            // [N][256]
  float* Bp = nullptr; int64_t Bp_cap = 0;            // [P][256]
  float* score = nullptr; int64_t score_cap = 0;      // [pair_cap]
  // training set
  int32_t *snode = nullptr, *sprim = nullptr, *sid = nullptr;
  float* slabel = nullptr;
  int64_t s_cap = 0, sid_cap = 0;
  float* loss = nullptr;
};

struct dp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  dp_system_desc desc{};
  dp_dev_sys hsys{};
  dp_dev_sys* d_sys = nullptr;
  int nbox = 1;  // footprint boxes per planar state

  // primitives
  int32_t P = 0, S = 0, maxlen = 0;
  std::vector<int32_t> h_off;
  std::vector<double> h_st;   // [S][nx] primitive states as uploaded (host search: trajectories)
  std::vector<double> h_ac;   // [S - P][nu] actions, or empty
  int32_t h_nu = 0;
  std::vector<double> sol_states, sol_actions;  // trajectory of the last dp_search
  int32_t sol_nu = 0;
  int32_t* d_off = nullptr;      // P+1
  double* d_st = nullptr;        // [nx][S]
  double* d_start = nullptr;     // [nx][P]
  double* d_last = nullptr;      // [nx][P] last states (applicability key of the reverse search)
  float* d_start32 = nullptr;    // [nx][P] FP32 copies of the keys (exact rejections with margins)
  float* d_last32 = nullptr;
  double* d_cost = nullptr;      // [P] motion cost (given, or ref_dt per action)
  double* d_geo = nullptr;       // [4*nbox][S] : relx rely c s (per box)
  double* d_sweep = nullptr;     // [4][P] canonical swept footprint box (planar) / centre box (quad)
  double* d_srange = nullptr;    // [2][nx][P] per-primitive min / max of each state component (canonical)
  double* d_act = nullptr;       // [A][nu] (quadrotor: nu = 4, 32-byte rows)
  float* d_act32 = nullptr;      // [A][nu] (16-byte rows)
  int32_t* d_goalidx = nullptr;  // [8][P]
  int32_t goal_num = -1;

  // world
  bool have_world = false;
  double world_lo[3] = {0, 0, 0}, world_hi[3] = {0, 0, 0};
  int32_t nb = 0, ns = 0;
  double* d_obs = nullptr;  // planar: [nb][4] cx cy hx hy ; 3-D: [nb][6] then [ns][4]

  // visited set (T_n)
  int64_t nv = 0, nv_cap = 0;
  double* d_vis = nullptr;         // [nx][nv_cap] append order
  double* d_vis_sorted = nullptr;  // [nx][nv_cap] sorted by cell
  int32_t* d_vis_cell = nullptr;   // scratch [nv_cap]
  int32_t* d_vis_perm = nullptr;   // [nv_cap] position in d_vis_sorted -> append index
  int32_t* d_vis_start = nullptr;  // [ncells+1]
  int64_t vis_cells_cap = 0;
  bool grid_dirty = true;
  dp_grid_geom grid{};
  double grid_radius = -1.0;

  // frontier nodes
  int32_t N = 0, N_cap = 0;
  double* d_nodes = nullptr;  // [nx][N_cap]
  double* d_gpar = nullptr;   // [N_cap]
  int32_t* d_app = nullptr;   // [N_cap+1] counts, then exclusive offsets
  uint64_t* d_app_mask = nullptr;  // [N][ceil(P/64)] applicable primitives per node
  int64_t app_mask_cap = 0;

  // pairs
  int64_t pair_cap = 0;
  int32_t* d_pnode = nullptr;
  int32_t* d_pprim = nullptr;
  int8_t* d_verdict = nullptr;
  int32_t* d_firstk = nullptr;
  int32_t* d_goal = nullptr;
  int8_t* d_res = nullptr;
  int8_t* d_flags = nullptr;  // bit0 dup-vs-visited, bit1 forced (goal), bit2 near-contact
  double* d_final = nullptr;  // [nx][pair_cap]
  double* d_g = nullptr;
  double* d_h = nullptr;

  // contender grid (batch duplicates)
  int32_t* d_cont_start = nullptr;  // [ncells+1]
  int32_t* d_cont_cursor = nullptr; // [ncells]
  int32_t* d_cont_idx = nullptr;    // [pair_cap]
  int64_t cont_start_cap = 0, cont_cursor_cap = 0;
  int32_t* d_node_novel = nullptr;  // [N_cap+1] NOVEL pairs per node, then offsets
  double* d_cont_x = nullptr;       // [nx][pair_cap] contender final states sorted by cell
  int16_t* d_conf_cnt = nullptr;    // [pair_cap] stored earlier conflicts per pair (-1: scan)
  int32_t* d_conf = nullptr;        // [pair_cap][DP_CONF_SLOTS] earlier conflicting contenders
  int32_t* d_conf_ov = nullptr;     // [pair_cap] offset of entries DP_CONF_SLOTS.. in the overflow pool
  int32_t* d_conf_pool = nullptr;   // [pool_cap] overflow entries (bump-allocated per batch)
  int64_t conf_pool_cap = 0;
  double round_r_dup = 0.0;         // duplicate radius of the batch in flight
  int32_t* d_tiles = nullptr;       // [tiles+1] partial sums of the large scan
  int64_t tiles_cap = 0;
  int32_t* d_node_cursor = nullptr; // [N_cap] first undecided pair per node
  int32_t* d_node_count = nullptr;  // [N_cap] NOVEL pairs decided per node
  long long* h_pinned = nullptr;    // pinned host words for device totals
  int32_t rounds = 32;              // resolution rounds launched per batch
  int32_t* d_chunk = nullptr;       // [N_cap+1] quadrotor: 64-pair chunks per node, scanned
  int32_t* d_nshort = nullptr;      // [N_cap] quadrotor: pairs whose primitive leaves the rate bounds
  int32_t* d_qperm = nullptr;       // [2][pair_cap] quadrotor: (pair, primitive) per node, rate-bound ones first
  int32_t* d_chunk_node = nullptr;  // quadrotor: node of each FP32 wave chunk
  int64_t chunk_node_cap = 0;
  int32_t* d_q64 = nullptr;         // [pair_cap] quadrotor: pairs the FP64 pass rolls out (DPC_Q64 of them)
  // quadrotor, per primitive state: attitude / rate trajectory and velocity
  // increments shared by every pair (k_quad_rot); per primitive: first state
  // whose attitude / rate leaves the bounds (FP64, FP32) or the FP32 margin
  double* d_rot64 = nullptr;        // [7][S]
  double* d_dv64 = nullptr;         // [S][4] (x y z, pad)
  float* d_dv32 = nullptr;          // [S][4]
  int32_t *d_rbad64 = nullptr, *d_rbad32 = nullptr, *d_rnear32 = nullptr;  // [P]
  double* d_qrec = nullptr;         // [P][16] quadrotor: the FP32 pass's 128-byte record per primitive
  double prim_reach = 0.0;          // max over primitives of max_k |p_k - p_0|_inf
  float* d_preach = nullptr;        // [P] quadrotor: max_k |p_k - p_0|_inf per primitive (rounded up)
  double reach_box[4] = {0, 0, 0, 0};  // planar: union of the canonical swept boxes (lo x, hi x, lo y, hi y)
  double reach_box_bwd[4] = {0, 0, 0, 0};  // the same relative to each primitive's last position, widened
  // settings of the other callers (tdbastar, dbrrt)
  int32_t* d_order = nullptr;       // [order_rows][P] pair order per node (NULL: ascending)
  int32_t order_rows = 0;
  int32_t forward = 1;              // 0: tdbastar's reverse search
  int32_t ncon = 0, con_cap = 0;    // tdbastar constraints
  double* d_con = nullptr;          // [con_cap] times, then [con_cap][nx] states
  double* d_targets = nullptr;      // [nx][N_cap] dbrrt targets of dp_expand_select
  int32_t targets_cap = 0;
  int32_t* d_sel_q = nullptr;       // [N_cap] chosen pair per node (-1: none)
  double* d_sel = nullptr;          // [N_cap][2 nx + 3] last state, first state, distance, g, h
  int32_t sel_cap = 0;
  bool sel_valid = false;           // the last expansion was a dp_expand_select
  bool select_mode = false;         // set while dp_expand_select enqueues its kernels
  double select_radius = 0.0, select_cost_jump = 0.0;
  int32_t select_first = 0;
  int16_t* d_node_obs = nullptr;    // [N_cap][DP_NODE_OBS] obstacles within a node's reach box
  int32_t* d_node_nobs = nullptr;   // [N_cap] their count (-1: more than DP_NODE_OBS)
  dp_mlp mlp;
  int32_t round_limit = 0;          // limit_branching_factor of the batch in flight
  dp_expand_params round_params{};  // parameters of the batch in flight

  // counters
  long long* d_counters = nullptr;  // [DPC_COUNT]
  long long* d_sticky = nullptr;    // [DPS_COUNT] never reset
  int32_t* d_wshift = nullptr;      // [2] quadrotor FP64 window growth (k_quad_window)
  int32_t* d_lneed = nullptr;       // [1] planar list window need (k_list_window)
  int64_t n_batches = 0;            // dp_expand calls
  int64_t host_completions = 0;     // batches finished by host-launched work in dp_get_counts
  // test-only limits (dp_debug_set_limits): inline conflict slots used
  // (<= DP_CONF_SLOTS), overflow pool capacity, FP64 window (0: defaults)
  int32_t conf_slots = DP_CONF_SLOTS;
  int64_t pool_limit = 0;
  int32_t fp64_window = 0;
  int32_t* d_total = nullptr;       // [1] number of pairs of the batch
  int32_t* d_scan_tmp = nullptr;    // [1024] block partials

  // timing
  // expansion stage events: 0 start, 4 applicability done, 1 pre-filter
  // done, 5 before the main check kernel, 6 after it, 2 check done, 7 lists
  // done, 3 end
  hipEvent_t ev[DP_EVENTS] = {};
  bool timing_valid = false;
  bool no_events = false;           // set while a multi-instance graph is captured
  bool defer_walk = false;          // set while a multi-instance batch launches: no walk (dp_walk_queue)
  int32_t poll_cap = 16;            // resolution: longest pause between polls (64-cycle sleeps)
  unsigned long long* d_trace = nullptr;  // [4 N_cap] resolution per node: start, end, waits, polls
  int64_t trace_cap = 0;
  dp_tile_config tiles{0, 2048, 4096, 256, 8192, 0.0f};  // launch shapes (dp_autotune)
  int64_t last_pairs = -1;
  void* radius = nullptr;           // device buffers of dp_visited_radius (k_radius.hip)
};

#define DP_HIP_RET(expr)                                                        \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      dp_set_error(std::string(#expr) + ": " + hipGetErrorString(e_));          \
      return DP_ERR_HIP;                                                        \
    }                                                                           \
  } while (0)

inline int dp_grid_blocks(int64_t work, int per_block, int max_blocks) {
  int64_t b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (int)b;
}

int dp_ensure_i32(int32_t** p, int64_t* cap, int64_t need);

// Records stage event k of the expansion on s (skipped while a
// multi-instance graph is captured).
inline int dp_record_event(dp_ctx* c, int k, hipStream_t s) {
  if (c->no_events) return DP_OK;
  DP_HIP_RET(hipEventRecord(c->ev[k], s));
  return DP_OK;
}
int dp_ensure_f64(double** p, int64_t* cap, int64_t need);
int dp_ensure_u64(uint64_t** p, int64_t* cap, int64_t need);

// ---- k_util.hip ---------------------------------------------------------
void dp_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
void dp_zero_counters(long long* c, hipStream_t s);
// Exclusive scan of data[0..n-1] into data[0..n]; data[n] = total.  Small n
// (one workgroup).  total (optional) receives min(total, cap); counters[slot]
// receives the total when slot >= 0.
void dp_scan_small(int32_t* data, int64_t n, int32_t* total, int64_t cap, long long* counters, int slot,
                   hipStream_t s);
// Same for large arrays: three passes over tiles of 4096 elements.  The
// length is *n_dev (at most n_max) when n_dev is set, else n_max.
int dp_scan_large(dp_ctx* c, int32_t* data, const int32_t* n_dev, int64_t n_max, long long* counters, int slot,
                  hipStream_t s);

// ---- k_grid.hip ---------------------------------------------------------
dp_grid_geom dp_make_grid(const dp_ctx* c, double r);
int dp_launch_visited_rebuild(dp_ctx* c, double radius, hipStream_t s);
int dp_launch_visited_append_novel(dp_ctx* c, int64_t novel_total, hipStream_t s);
int dp_launch_resolve(dp_ctx* c, const dp_expand_params* p, double r_dup, hipStream_t s);
int dp_launch_more_rounds(dp_ctx* c, int rounds, hipStream_t s);
// Resolution from scratch after a window overflow: every FREE pair
// undecided again, resolution counters cleared, lists and walk rebuilt.
int dp_launch_resolve_again(dp_ctx* c, hipStream_t s);
// The ordered walks of a multi-instance batch as one work queue: contexts
// expanded with defer_walk set stop before their walk, and one launch per
// system class walks every instance, interleaved, stamping each instance's
// completion clock into stamps[i].  prepare (host, synchronous: it first
// waits for `prev`, the stream of the previous launch, if the arrays change)
// uploads the contexts' walk arguments; it runs after their pipelines were
// launched and before any capture of the queue launch.
struct dp_walk_queue;
dp_walk_queue* dp_walk_queue_create(int32_t n);
void dp_walk_queue_destroy(dp_walk_queue* q);
int dp_walk_queue_prepare(dp_walk_queue* q, dp_ctx* const* ctx, hipStream_t prev);
int dp_walk_queue_launch(dp_walk_queue* q, unsigned long long* stamps, hipStream_t s);

// ---- k_radius.hip -------------------------------------------------------
int dp_noalloc_init_device(int device);
void dp_radius_free(dp_ctx* c);

// ---- k_quad.hip ---------------------------------------------------------
int dp_launch_quad(dp_ctx* c, const dp_expand_params* p, double goal_bound, double r_dup, hipStream_t s);
// FP64 rollout of the pairs deferred past their node's FP64 window (after a
// window overflow), then their visited-set duplicate test.
int dp_launch_quad_deferred(dp_ctx* c, hipStream_t s);
// Per-primitive attitude / rate trajectories (after a primitive upload).
int dp_launch_quad_rot(dp_ctx* c, hipStream_t s);
// FP64 window of the batch in FREE pairs (INT32_MAX: none).
int32_t dp_fp64_window(const dp_ctx* c, int32_t limit);

// ---- k_mlp.hip ----------------------------------------------------------
int dp_mlp_alloc(dp_ctx* c, int32_t P, uint64_t seed);
int64_t dp_mlp_param_count();
int dp_mlp_fwd_bwd(dp_ctx* c, const uint16_t* X, const float* y, int32_t B, const int32_t* sample,
                   const int32_t* sprim, float* loss, hipStream_t s);
// Score the pairs of the batch (after k_app fill); rejects below the
// threshold when `reject`.  Pairs left for the exact check get verdict -1.
int dp_launch_surrogate(dp_ctx* c, bool reject, hipStream_t s);
void dp_mlp_free(dp_ctx* c);

// ---- k_expand.hip -------------------------------------------------------
int dp_launch_expand(dp_ctx* c, const dp_expand_params* p, hipStream_t s);
int dp_launch_count(dp_ctx* c, const dp_expand_params* p, hipStream_t s);
// Bumped by every device (re)allocation and every change of state a captured
// multi-instance graph bakes in; a graph captured at an older epoch is
// captured again before it is replayed.
void dp_bump_epoch();
uint64_t dp_epoch();
int dp_prepare_goal_index(dp_ctx* c, const dp_expand_params* p, hipStream_t s);
int dp_size_pairs(dp_ctx* c, const dp_expand_params* p, hipStream_t s);

// Any visited state within r of x in run k (0..8) of x's 27-cell block?  The
// block is 9 runs of 3 cells adjacent in memory (x is the fastest grid axis):
// run k covers rows c1 + k%3 - 1 and layers c2 + k/3 - 1 (the rules of
// dp_cell_neighbor), x from c0 - 1 to c0 + 1 clipped to the grid.
template <int NX>
__device__ __forceinline__ bool dp_grid_run_any(const dp_dev_sys& S, const dp_grid_geom& G, const int32_t* start,
                                             const double* pts, int64_t stride, const double* x, int k, double r,
                                             double T) {
  int32_t c[3];
  dp_cell_coords(G, x, c);
  const int32_t n1 = c[1] + k % 3 - 1;
  const int32_t d2 = k / 3 - 1;
  if (n1 < 0 || n1 >= G.dims[1]) return false;
  int32_t n2 = c[2] + d2;
  if (G.dims[2] == 1) {
    if (d2 != 0) return false;
    n2 = 0;
  } else if (G.periodic2) {
    n2 = n2 < 0 ? n2 + G.dims[2] : (n2 >= G.dims[2] ? n2 - G.dims[2] : n2);
  } else if (n2 < 0 || n2 >= G.dims[2]) {
    return false;
  }
  const int64_t row = ((int64_t)n2 * G.dims[1] + n1) * G.dims[0];
  const int32_t x0 = c[0] > 0 ? c[0] - 1 : 0, x1 = c[0] + 1 < G.dims[0] ? c[0] + 1 : G.dims[0] - 1;
  const int32_t e = start[row + x1 + 1];
  for (int32_t p = start[row + x0]; p < e; ++p)
    if (dp_within_col_pf<NX>(S, x, pts, stride, p, r, T)) return true;
  return false;
}
