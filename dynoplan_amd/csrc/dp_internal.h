// dp_internal.h -- context layout shared by the C-ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "dp_device.h"

// Largest obstacle count staged in LDS by one workgroup.
#define DP_MAX_OBS 1024
// Per-pair resolution value while undecided (ordered duplicate resolution).
#define DP_R_UNDECIDED 0x7f
// Pad added to swept boxes of the broad phase (m): bounds the rounding of
// "canonical sweep + offset" versus per-state "state + offset" by 1e6x.
#define DP_SWEEP_PAD 1e-9
// Bound on spin iterations of one wave in the ordered resolution; reaching
// it sets an error flag instead of waiting forever.
#define DP_SPIN_LIMIT (1 << 24)

void dp_set_error(const std::string& msg);

// Device-side counters, one slot per dp_counts field (int64 each) plus flags.
enum {
  DPC_PAIRS = 0, DPC_BOUNDS, DPC_COLLIDE, DPC_FREE, DPC_SURR, DPC_NOVEL, DPC_DUPV, DPC_DUPB,
  DPC_CAPPED, DPC_GOAL, DPC_CHECKED, DPC_RECHECK, DPC_SPINS, DPC_ERRFLAG, DPC_OVERFLOW,
  DPC_TICKET, DPC_COUNT
};

// Uniform grid over the position components (2 or 3), cell = r/w0 padded.
struct dp_grid_geom {
  double lo[3];
  double inv_cell;  // 1 / cell
  int32_t dims[3];
  int32_t pos_dim;
  int64_t ncells;
};

__host__ __device__ inline int32_t dp_cell_axis(double x, double lo, double inv_cell, int32_t dim) {
  double f = (x - lo) * inv_cell;
  int32_t c = f < 0.0 ? 0 : (f >= (double)(dim - 1) ? dim - 1 : (int32_t)f);
  return c;
}

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
  int32_t* d_off = nullptr;      // P+1
  double* d_st = nullptr;        // [nx][S]
  double* d_start = nullptr;     // [nx][P]
  double* d_geo = nullptr;       // [4*nbox][S] : relx rely c s (per box)
  double* d_sweep = nullptr;     // [4][P] canonical swept footprint box (planar) / centre box (quad)
  double* d_act = nullptr;       // [nu][A]
  float* d_act32 = nullptr;      // [nu][A]
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

  // pairs
  int64_t pair_cap = 0;
  int32_t* d_pnode = nullptr;
  int32_t* d_pprim = nullptr;
  int8_t* d_verdict = nullptr;
  int32_t* d_firstk = nullptr;
  int32_t* d_goal = nullptr;
  int32_t* d_res = nullptr;
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

  // counters
  long long* d_counters = nullptr;  // [DPC_COUNT]
  int32_t* d_total = nullptr;       // [1] number of pairs of the batch
  int32_t* d_scan_tmp = nullptr;    // [1024] block partials

  // timing
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  bool timing_valid = false;
  int64_t last_pairs = -1;
};

// ---- launchers (k_expand.hip) -------------------------------------------
int dp_launch_expand(dp_ctx* c, const dp_expand_params* p, hipStream_t s);
int dp_launch_visited_rebuild(dp_ctx* c, double radius, hipStream_t s);
int dp_launch_visited_append_novel(dp_ctx* c, int64_t novel_total, hipStream_t s);
