/*
 * dynoplan_amd.h -- C-ABI of the MI355X db-A* node-expansion path.
 *
 * One call of dp_expand() does, for a batch of N frontier nodes, what the
 * reference's inner expansion loop does for one node at a time
 * (src/dbastar/dbastar.cpp:806-919):
 *
 *   canonical_state / offset           dbastar.hpp:273-274   (Model_robot)
 *   applicability query  nearestR(c, alpha*delta) over primitive starts
 *                                       dbastar.hpp:281, dbastar.cpp:585-586
 *   transform_primitive + bounds        dbastar.cpp:402-446, dbastar.hpp:166-215
 *   collision check (first index)       dbastar.cpp:459-489
 *   check_goal                          dbastar.cpp:492-512, 825-829
 *   g / h                               dbastar.cpp:833-847
 *   duplicate test + insert             dbastar.cpp:850-881
 *   branching cap                       dbastar.cpp:916-918
 *
 * Batch semantics (the canonical order): pairs (node i, primitive j) are
 * visited in ascending (i, j) order.  Every applicable pair receives a
 * verdict (FREE / BOUNDS / COLLIDE / SURROGATE_REJECT) before any cap.  The
 * duplicate decision of a FREE pair is taken against the visited set plus
 * every NOVEL pair that precedes it in canonical order, exactly as the
 * reference inserts each novel node into T_n at once (dbastar.cpp:875-876).
 * The reference leaves a node's loop after a FREE pair once its NOVEL count
 * reaches limit_branching_factor (dbastar.cpp:916-918): every later FREE
 * pair of the node is CAPPED, and, as the reference never reaches it, has no
 * goal index, final state, g or h (dp_fetch_pairs gives -1 / NaN).  The
 * reference shuffles the applicable list (dbastar.hpp:302-303); the
 * canonical order is the identity shuffle.
 *
 * Conventions: every entry point returns DP_OK (0) or a negative DP_ERR_*;
 * dp_last_error() gives a thread-local message.  No exceptions cross the
 * boundary.  A context is used by one host thread at a time; work is issued
 * on the stream passed in (NULL = the context's own stream).  Device memory
 * is owned by the context; host buffers are owned by the caller.
 */
#ifndef DYNOPLAN_AMD_H
#define DYNOPLAN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  DP_OK = 0,
  DP_ERR_ARG = -1,      /* invalid argument (sizes, NULL pointers, alpha not in (0,1), ...) */
  DP_ERR_HIP = -2,      /* a HIP runtime call failed */
  DP_ERR_STATE = -3,    /* call out of order (no primitives / world uploaded, ...) */
  DP_ERR_CAPACITY = -4, /* output capacity exceeded */
  DP_ERR_IO = -5,       /* file could not be read */
  DP_ERR_FORMAT = -6,   /* file content malformed */
  DP_ERR_NOMEM = -7,    /* allocation failed */
  DP_ERR_NOTCONV = -8   /* a bounded device loop did not finish (reported, never hangs) */
};

/* ---- systems (dynobench::Model_robot instances on the path) ------------ */
enum dp_system_id {
  DP_SYS_UNICYCLE1 = 0,   /* unicycle1_v0: x y theta | v w                     */
  DP_SYS_UNICYCLE2 = 1,   /* unicycle2_v0: x y theta v w | a alpha             */
  DP_SYS_QUADROTOR = 2,   /* quad3d / quadrotor_v0: p(3) q(xyzw) v(3) w(3) | 4 rotors */
  DP_SYS_CAR_TRAILER = 3  /* car_first_order_with_1_trailers: x y th0 th1 | v phi */
};

#define DP_MAX_NX 16

/* Per-pair verdict (pre-cap).  DP_V_CONSTRAINT: valid but for a tdbastar
 * constraint (dp_set_constraints). */
enum { DP_V_FREE = 0, DP_V_BOUNDS = 1, DP_V_COLLIDE = 2, DP_V_SURROGATE_REJECT = 3, DP_V_CONSTRAINT = 4 };
/* Duplicate/insert resolution of a FREE pair. */
enum { DP_R_NONE = 0, DP_R_NOVEL = 1, DP_R_DUP_VISITED = 2, DP_R_DUP_BATCH = 3, DP_R_CAPPED = 4 };

/* System description: the fields of dynobench::Model_robot the path reads
 * (nx, nu, ref_dt, x_lb/x_ub for is_state_valid, distance_weights,
 * lower_bound_time speeds, collision geometry).  dp_system_default() fills
 * the values this library uses for each system; position bounds are taken
 * from the world (robot_factory(..., p_lb, p_ub) in the reference). */
typedef struct dp_system_desc {
  int32_t system;              /* dp_system_id */
  int32_t nx, nu;
  int32_t pos_dim;             /* 2 (planar) or 3 */
  double dt;                   /* ref_dt */
  double x_lb[DP_MAX_NX];      /* is_state_valid lower bounds (position rows overwritten by the world) */
  double x_ub[DP_MAX_NX];
  double dist_w[4];            /* distance_weights, one per metric component group */
  double max_vel, max_angular_vel, max_acc, max_angular_acc; /* lower_bound_time */
  double body_half[2];         /* planar: robot box half length/width; quadrotor: [0] = sphere radius */
  double trailer_half[2];      /* car+trailer: trailer box half length/width */
  double hitch;                /* car+trailer: hitch length; trailer centre at p - hitch*(cos th1, sin th1) */
  double wheelbase;            /* car+trailer: tractor wheelbase */
  double quad[8];              /* quadrotor: mass, g, arm, t2t, Jx, Jy, Jz, -          */
  double near_eps;             /* FP32 path: margin under which a verdict is re-checked in FP64: near_eps
                                  (m) for clearances and position bounds, 10 x for velocity and 100 x for
                                  angular-rate bounds */
} dp_system_desc;

/* Motion primitives (dynoplan::Motion, ompl/robots.h:206-235), row-major as
 * loaded from file.  Primitive j owns states [offsets[j], offsets[j+1]) and
 * actions [offsets[j]-j, offsets[j+1]-j-1). */
typedef struct dp_prims {
  int32_t P;
  int32_t nx, nu;
  int32_t reserved;
  const int32_t* offsets;      /* P+1 */
  const double* states;        /* [offsets[P]][nx] */
  const double* actions;       /* [offsets[P]-P][nu] (may be NULL for translate-only systems) */
  const double* cost;          /* [P] or NULL (then ref_dt * #actions, robots.cpp:3153-3159) */
} dp_prims;

/* Environment (dynobench env: bounds + obstacles). */
typedef struct dp_world {
  int32_t n_boxes, n_spheres;
  double lo[3], hi[3];         /* environment min/max; position bounds of is_state_valid */
  const double* boxes;         /* [n_boxes][6]  cx cy cz hx hy hz (planar: z ignored) */
  const double* spheres;       /* [n_spheres][4] cx cy cz r */
} dp_world;

/* Options_dbastar fields the expansion reads (dbastar/options.hpp:105-156). */
typedef struct dp_expand_params {
  float delta;                     /* discontinuity bound (float, as the reference) */
  float alpha;                     /* in (0,1) */
  double delta_factor_goal;        /* goal radius = delta_factor_goal * delta */
  double cost_delta_factor;
  int32_t limit_branching_factor;  /* the reference's break (dbastar.cpp:916-918), a size_t NOVEL count
                                      compared with this int: 0 decides only the first FREE pair of a
                                      node, a negative value never caps */
  int32_t num_check_goal;          /* intermediate goal checks (reference: 4) */
  int32_t check_intermediate_goal; /* reference: true */
  int32_t use_surrogate;           /* learned pre-filter (needs a trained surrogate) */
  int32_t fp32_rollout;            /* quadrotor: FP32 rollout + FP64 near-contact re-check */
  int32_t reserved;
  double goal[DP_MAX_NX];
} dp_expand_params;

/* Per-batch counters.  "pairs" is the unit of the throughput metric: one
 * (frontier node, applicable primitive) pair that received a verdict. */
typedef struct dp_counts {
  int64_t nodes, pairs;
  int64_t bounds, collide, free_, surrogate_reject;
  int64_t novel, dup_visited, dup_batch, capped;
  int64_t goal_hits;        /* FREE pairs with chosen_index != -1 */
  int64_t checked;          /* pairs that reached the collision check (num_col_motions) */
  int64_t recheck_fp64;     /* FP32 pairs re-evaluated in FP64 near contact */
  int64_t resolve_rounds;   /* diagnostic: rounds of the batch duplicate resolution */
  int64_t conf_pool_used;   /* diagnostic: earlier-conflict entries past the inline slots */
  int64_t fp64_deferred;    /* quadrotor: FREE pairs past their node's FP64 window (rolled out in FP64
                               only if the node's cap did not fall before them) */
  int64_t constraint;       /* pairs valid but for a tdbastar constraint (DP_V_CONSTRAINT) */
} dp_counts;

/* Caller-owned host buffers for dp_fetch_pairs (any pointer may be NULL). */
typedef struct dp_pairs_out {
  int64_t capacity;
  int32_t* node;
  int32_t* prim;
  int8_t* verdict;          /* DP_V_* */
  int32_t* first_k;         /* BOUNDS/COLLIDE: first offending state index, else -1 */
  int32_t* goal_index;      /* chosen_index of check_goal, -1 if none */
  int8_t* resolution;       /* DP_R_* */
  double* final_state;      /* [capacity][nx]: state stored in the new node (FREE pairs) */
  double* g;                /* FREE pairs: gScore */
  double* h;                /* FREE pairs: hScore */
} dp_pairs_out;

/* ---- entry points ------------------------------------------------------ */
typedef struct dp_ctx dp_ctx;

/* Thread-local description of the last error. */
const char* dp_last_error(void);

/* Defaults for a system (replaces dynobench::robot_factory + model YAML). */
int dp_system_default(int32_t system, dp_system_desc* out);

/* Context on HIP device `device`.  Returns NULL on failure (see dp_last_error). */
dp_ctx* dp_create(const dp_system_desc* sys, int32_t device);
void dp_destroy(dp_ctx* ctx);

/* Upload primitives (replaces T_m->add over motions, dbastar.cpp:564-575,
 * and the per-motion fcl shapes of compute_col_shape, robots.cpp:3190-3215).
 * Only the first max_motions are used (<= 0: all). */
int dp_upload_primitives(dp_ctx* ctx, const dp_prims* prims, int32_t max_motions);

/* Upload the environment (load_env; obstacles become LDS-staged geometry). */
int dp_upload_world(dp_ctx* ctx, const dp_world* world);

/* Visited-node set for the duplicate test (T_n, dbastar.cpp:578-579). */
int dp_visited_reset(dp_ctx* ctx);
int dp_visited_append(dp_ctx* ctx, const double* states /*[n][nx]*/, int64_t n);
/* Append the NOVEL pairs of the last dp_expand() (T_n->add, dbastar.cpp:875-876). */
int dp_visited_append_novel(dp_ctx* ctx);
int64_t dp_visited_size(const dp_ctx* ctx);

/* Radius query over the visited set (T_n->nearestR, nigh_custom_spaces.hpp:236-247,
 * the rewiring query of dbastar.cpp:883-912), on the device grid: for each
 * query state, the indices (append order) of visited states with distance
 * <= radius, ascending, CSR output.  out_offsets has n+1 entries (always
 * complete); out_idx receives the first min(total, idx_cap) entries and the
 * call returns DP_ERR_CAPACITY when total > idx_cap.  Host buffers;
 * synchronous; device buffers grow on demand. */
int dp_visited_radius(dp_ctx* ctx, const double* queries /*[n][nx]*/, int64_t n, double radius,
                      int64_t* out_offsets, int64_t* out_idx, int64_t idx_cap);

/* dp_visited_radius without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * that must not call hipMalloc, tight serving loops).  Queries, offsets and
 * indices are DEVICE buffers of the caller ([n][nx] doubles, n+1 and idx_cap
 * int64).  Its per-query counts live in a fixed per-device scratch array
 * sized for DP_NOALLOC_MAX_QUERIES queries; larger n is rejected
 * (DP_ERR_ARG).  The visited grid must already be built for a radius >=
 * `radius` (by dp_visited_radius or dp_expand after the last change of the
 * visited set), else DP_ERR_STATE.  Work is enqueued on `stream` (NULL = the
 * context stream) and the call waits for it.  DP_ERR_CAPACITY when the
 * total exceeds idx_cap (entries beyond it are not written) or a query has
 * more than 4096 neighbours (its list is then not sorted).  All calls on one
 * device share the scratch array: calls that overlap on one device race. */
#define DP_NOALLOC_MAX_QUERIES 65536
int dp_visited_radius_noalloc(dp_ctx* ctx, const double* d_queries, int64_t n, double radius,
                              int64_t* d_offsets, int64_t* d_idx, int64_t idx_cap, void* stream);

/* Pre-size device buffers (optional; dp_expand grows them on demand). */
int dp_reserve(dp_ctx* ctx, int32_t max_nodes, int64_t max_pairs);

/* Copy N frontier nodes (row-major [N][nx]) and their gScore to the device. */
int dp_upload_nodes(dp_ctx* ctx, const double* nodes, const double* g_parent, int32_t N);

/* Expand the uploaded nodes: enqueue all kernels on `stream` (hipStream_t,
 * NULL = context stream) and return without synchronising. */
int dp_expand(dp_ctx* ctx, const dp_expand_params* params, void* stream);

/* dp_upload_nodes + dp_expand + synchronise: the batched counterpart of one
 * pass of the reference's expansion loop. */
int dp_expand_batch(dp_ctx* ctx, const double* nodes, const double* g_parent, int32_t N,
                    const dp_expand_params* params, void* stream);

/* Counters of the last dp_expand (synchronises the stream).  Finishes the
 * batch first if its device pipeline left work (a resolution wave that ran
 * out of its wait bound, or a quadrotor node whose cap fell past its FP64
 * window): dp_health counts such batches. */
int dp_get_counts(dp_ctx* ctx, dp_counts* out);

/* Totals over the context's life, never reset (no counterpart in the
 * reference): expansions issued, resolution waves that gave up waiting
 * (their nodes finished by host-launched rounds), FP64 window overflows,
 * and batches completed by host-launched work in dp_get_counts.  A caller
 * timing dp_expand calls checks that the last three did not move. */
typedef struct dp_health_info {
  int64_t batches;
  int64_t wave_giveups;
  int64_t window_overflows;
  int64_t host_completions;
} dp_health_info;
int dp_health(dp_ctx* ctx, dp_health_info* out);

/* Test-only limits (no counterpart in the reference) that force the rare
 * paths of the resolution: inline earlier-conflict slots used per pair
 * (1..64; 0 restores 64), overflow pool entries (0: the full pool; a list
 * that does not fit is scanned from the grid when reached) and the window in
 * FREE pairs per node: the quadrotor's FP64 window (0: 3 * limit + 32,
 * doubled for a context's later batches each time one of its batches
 * overflowed it, up to 2^10 times) and, for planar systems, the pairs that
 * get stored conflict lists (0: 3 * limit + 32; a pair past it is scanned
 * from the grid when reached).  Results do not depend on them. */
int dp_debug_set_limits(dp_ctx* ctx, int32_t conf_slots, int64_t pool_entries, int32_t fp64_window);

/* Copy the per-pair results of the last dp_expand into caller buffers, in
 * canonical (node, primitive) order.  Returns DP_ERR_CAPACITY if the batch
 * has more pairs than out->capacity. */
int dp_fetch_pairs(dp_ctx* ctx, dp_pairs_out* out);

/* The stream the context issues its own work on (hipStream_t). */
void* dp_stream(dp_ctx* ctx);

/* Device time (ms) of the kernels of the last dp_expand, split by stage
 * (HIP events on the expansion stream), ms5[5]: 0 applicability, 1 check,
 * 2 duplicate resolution, 3 total, 4 learned pre-filter (0 without it). */
int dp_last_timing(dp_ctx* ctx, float* ms5);
/* The same split per kernel group, ms7[7]: 0 applicability, 1 learned
 * pre-filter, 2 before the main check kernel (node obstacle lists; quadrotor
 * chunk offsets), 3 the main check kernel (k_check_planar; quadrotor
 * k_quad32), 4 after it (quadrotor: FP64 window, FP64 pass, visited
 * duplicates), 5 contender grid and conflict lists, 6 the ordered walk. */
int dp_last_timing_detail(dp_ctx* ctx, float* ms7);

/* Launch shapes of the check stage.  Every configuration gives identical
 * results; dp_autotune picks the fastest for the uploaded batch. */
typedef struct dp_tile_config {
  int32_t lanes_per_pair;   /* planar check: lanes per pair, 2 to 64 (0: from the longest primitive;
                                 a shape whose LDS would not fit widens) */
  int32_t check_blocks;     /* planar check: workgroups (grid-stride) */
  int32_t quad32_blocks;    /* quadrotor FP32 pass: workgroups */
  int32_t quad64_threads;   /* quadrotor FP64 pass: threads per workgroup, 64, 128 or 256 */
  int32_t quad64_blocks;    /* quadrotor FP64 pass: workgroups */
  float check_ms;           /* dp_autotune: check-stage time of the chosen shape */
} dp_tile_config;

int dp_set_tiles(dp_ctx* ctx, const dp_tile_config* tiles);
int dp_get_tiles(const dp_ctx* ctx, dp_tile_config* tiles);
/* Times the check stage of the uploaded nodes under each candidate shape
 * (median of 3 after a warm-up), keeps the fastest and returns it; `trials`
 * (optional, up to 64) receives every candidate with its time. */
int dp_autotune(dp_ctx* ctx, const dp_expand_params* params, dp_tile_config* best, dp_tile_config* trials,
                int32_t* n_trials);

/* ---- primitive files (load_motion_primitives_new, src/ompl/robots.cpp:3061-3174) ---- */
enum { DP_FMT_AUTO = 0, DP_FMT_YAML = 1, DP_FMT_JSON = 2, DP_FMT_MSGPACK = 3, DP_FMT_BOOST = 4 };

/* Host arrays of a loaded primitive set, in the dp_prims layout (allocated
 * by dp_load_primitives, released by dp_free_primitives). */
typedef struct dp_prim_set {
  int32_t P, nx, nu, reserved;
  int32_t* offsets;            /* P+1 */
  double* states;              /* [offsets[P]][nx] */
  double* actions;             /* [offsets[P]-P][nu] */
  double* cost;                /* [P] */
} dp_prim_set;

/* Load a primitive file as load_motion_primitives_new does: format from the
 * extension when DP_FMT_AUTO (robots.cpp:3070-3080), truncation to
 * max_motions (<= 0: all, :3104-3105), first/last state + noise*U[-1,1]
 * drawn from a seeded stream (the reference adds 1e-7 from std::rand,
 * :3113-3127; noise = 0 disables it), quaternion normalisation for the
 * quadrotor (:3130-3137), cost = ref_dt * #actions when the first cost
 * exceeds 1e5 (:3153-3159).  Primitive j gets idx = j (:3170-3173). */
int dp_load_primitives(const char* path, int32_t format, int32_t system, int32_t max_motions, double noise,
                       uint64_t seed, dp_prim_set* out);
void dp_free_primitives(dp_prim_set* set);

/* Write a primitive set as Boost binary (.bin), msgpack or YAML; the
 * readers above load these files back bit-exactly (format conversion of
 * main_primitives' bintoyaml / yamltobin modes, main_primitives.cpp:4-20). */
int dp_save_primitives(const char* path, int32_t format, const dp_prim_set* set);

/* ---- problem files (dynobench problem YAML; benchmark.py:330-350) ------------
 * environment: {min, max, obstacles: [{type: box, center, size (full side
 * lengths)} | {type: sphere, center, size: [radius]}]}, robots: [{type,
 * start, goal}] (the first robot is read).  `world` is ready for
 * dp_upload_world (boxes as centre + half sizes); dp_free_problem releases
 * its arrays. */
typedef struct dp_problem {
  int32_t system;                  /* DP_SYS_* from the robot type's prefix */
  int32_t nx;
  char robot_type[64];             /* e.g. unicycle1_v0 */
  double start[DP_MAX_NX], goal[DP_MAX_NX];
  double lo[3], hi[3];             /* environment min / max */
  int32_t n_boxes, n_spheres;
  double* boxes;                   /* [n_boxes][6] cx cy cz hx hy hz */
  double* spheres;                 /* [n_spheres][4] cx cy cz r */
} dp_problem;

int dp_load_problem(const char* path, dp_problem* out);
void dp_free_problem(dp_problem* p);
int dp_save_problem(const char* path, const dp_problem* p);

/* ---- learned collision-feasibility surrogate --------------------------------
 * No counterpart in the reference.  An MLP over a 32x32 occupancy patch
 * (0.1 m cells) around the node and a learned 64-d primitive embedding
 * (3 hidden layers of 256, BF16 MFMA with FP32 accumulation) scores each
 * pair; with dp_expand_params.use_surrogate = 1, pairs scoring below the
 * calibrated threshold get DP_V_SURROGATE_REJECT and skip the exact check,
 * every other pair is checked exactly.  Planar systems. */
typedef struct dp_surrogate_stats {
  float threshold;
  float reserved;
  int64_t pairs, feasible, infeasible;
  int64_t rejectable;            /* infeasible pairs scoring below the threshold */
  int64_t colliding, rejectable_colliding;
} dp_surrogate_stats;

int dp_surrogate_init(dp_ctx* ctx, uint64_t seed);
/* Exact-checker labels (FREE = 1) of the pairs of `nodes`, then `epochs`
 * passes of minibatch Adam; loss_out[epochs] (optional) gets the epoch losses. */
int dp_surrogate_train(dp_ctx* ctx, const double* nodes, int32_t N, const dp_expand_params* params, int32_t epochs,
                       int32_t batch, float lr, uint64_t seed, float* loss_out);
/* Threshold from the scores of the FREE pairs of held-out `nodes`: their
 * lowest score s1 minus max(margin, tail * (s10 - s1)), s10 the tenth lowest
 * (the spread of the lower tail widens the margin where the tail is long). */
int dp_surrogate_calibrate(dp_ctx* ctx, const double* nodes, int32_t N, const dp_expand_params* params,
                           float margin, float tail, dp_surrogate_stats* out);
int dp_surrogate_set_threshold(dp_ctx* ctx, float threshold);
/* Scores of the last expansion run with use_surrogate != 0. */
int dp_surrogate_scores(dp_ctx* ctx, float* out, int64_t n);
int64_t dp_surrogate_param_count(void);
/* Forward + backward on explicit inputs X [B][1088] (patch | embedding),
 * labels y [B], parameters `params`: mean BCE loss and its gradient. */
int dp_surrogate_grad(dp_ctx* ctx, const float* params, const float* X, const float* y, int32_t B, float* loss,
                      float* grad);
/* Current FP32 master parameters (dp_surrogate_param_count() floats, the
 * layout W1[256][1088] b1 W2[256][256] b2 W3[256][256] b3 w4[256] b4) and
 * primitive embeddings [P][64]; either pointer may be NULL. */
int dp_surrogate_get(dp_ctx* ctx, float* params, float* emb);
/* Numerics check of the training GEMM: C = A B, inputs rounded to BF16. */
int dp_debug_gemm_bf16(dp_ctx* ctx, const float* A, const float* B, int32_t M, int32_t N, int32_t K, float* C);

/* Quadrotor rollout error report, every state: the states of the listed
 * pairs (indices in the last dp_expand's canonical pair order) as the FP32
 * pass forms them (bits = 32, positions made absolute in FP64) or as the
 * FP64 pass forms them (bits = 64), into out[n][L][13] with L the longest
 * primitive's state count (rows past a pair's own length untouched).  No
 * counterpart in the reference. */
int dp_debug_rollout_states(dp_ctx* ctx, const int32_t* pairs, int32_t n, int32_t bits, double* out);

/* Diagnostics of the duplicate resolution (no counterpart in the reference):
 * enable > 0 records, for each node of later batches, its wave's timeline;
 * enable < 0 stops.  With out != NULL, copies the record of the last batch,
 * out[8n + 1]: per node (start clock, end clock, cycles waiting on other
 * nodes' verdicts, cycles evaluating conflict lists, cycles walking chunks,
 * chunks, waits, polls), then the clock rate in kHz. */
int dp_debug_resolve_trace(dp_ctx* ctx, int32_t enable, uint64_t* out, int64_t n);

/* ---- db-A* search on the expansion path (dbastar.cpp:523-1010) -------------
 * The host loop of dbastar(): open list ordered by float fScore then higher
 * gScore, goal test on every popped node, expansion of the popped nodes with
 * dp_expand_batch, a node for every NOVEL pair, rewiring of the neighbours
 * of duplicate pairs, the solution path.  Uses the context's primitives and
 * world, and replaces its visited set with the search's nodes. */
enum { DP_SEARCH_SOLVED = 0, DP_SEARCH_MAX_EXPANDS = 1, DP_SEARCH_EMPTY_QUEUE = 2, DP_SEARCH_MAX_TIME = 3 };

typedef struct dp_search_options {  /* Options_dbastar (options.hpp:105-156) */
  float delta;                      /* .3 */
  float alpha;                      /* .5 */
  double delta_factor_goal;         /* 1 */
  double cost_delta_factor;         /* 1 */
  int64_t max_expands;              /* 1e6 */
  double search_timelimit_ms;       /* 1e4 */
  int32_t limit_branching_factor;   /* 20 */
  int32_t heuristic;                /* 0: lower_bound_time to the goal, -1: blind */
  int32_t batch_nodes;              /* nodes popped per expansion; 1 = the reference's order */
  int32_t use_surrogate;
  int32_t fp32_rollout;
  int32_t reserved;
} dp_search_options;

typedef struct dp_search_info {     /* Out_info_db and Time_benchmark fields */
  int32_t solved, status;           /* DP_SEARCH_* */
  double cost;                      /* gScore of the solution (or closest) node */
  double distance_to_goal;
  int64_t expands, all_nodes, closed, rewires, gpu_batches;
  int64_t pairs, checked;           /* checked = num_col_motions */
  int64_t traj_states, traj_actions;
  double time_search_ms, time_gpu_ms;
} dp_search_info;

void dp_search_options_default(dp_search_options* opt);
int dp_search(dp_ctx* ctx, const double* start, const double* goal, const dp_search_options* opt,
              dp_search_info* info);
/* The last search's trajectory: states [traj_states][nx], actions
 * [traj_actions][nu] (either pointer may be NULL). */
int dp_search_trajectory(dp_ctx* ctx, double* states, double* actions);
/* Sizes of that trajectory (after dp_search or dp_idbastar). */
int dp_search_trajectory_size(const dp_ctx* ctx, int64_t* n_states, int64_t* n_actions);

/* ---- iDb-A* schedule (src/idbastar/idbastar.cpp:128-312) ---------------------
 * Repeated dp_search runs with delta and the number of primitives scheduled
 * as the reference's new_schedule: iteration 0 uses delta_0 and
 * num_primitives_0; later iterations multiply delta by delta_rate after a
 * solved run (by 0.9999 otherwise) and the primitive count by
 * num_primitives_rate, rounding to float delta and truncating to an integer
 * count each time (Options_dbastar's float delta and size_t max_motions).
 * After a solved run the solution trajectory is cut into `cuts` pieces of
 * floor(k * A / cuts) .. floor((k+1) * A / cuts) of its A actions, each made
 * canonical (the offset dims of its first state subtracted from every
 * state), its first and last states perturbed by 1e-7 U[-1,1] (one seeded
 * splitmix64 stream, first state then last, pieces in order; quadrotor
 * quaternions renormalised), and the pieces are inserted before every
 * earlier primitive in cut order (idbastar.cpp:234-291); then the
 * primitives are uploaded again.  Trajectory optimisation between the
 * iterations (crocoddyl in the reference) is not part of this library: the
 * raw search solution stands in for the optimised one. */
enum { DP_IDB_MAX_IT = 0, DP_IDB_TIME_LIMIT = 1, DP_IDB_MAX_SOLUTION = 2 };  /* EXIT_CRITERIA */

typedef struct dp_idbastar_options {  /* Options_idbAStar (idbastar.hpp:45-60) */
  double delta_0;                     /* .3 */
  int64_t num_primitives_0;           /* 1000 */
  double delta_rate;                  /* .9 */
  double num_primitives_rate;         /* 1.5 */
  double timelimit_s;                 /* 10 */
  int32_t max_it;                     /* 10 */
  int32_t max_num_sol;                /* 5 */
  int32_t add_primitives;             /* 1 */
  int32_t cuts;                       /* 5 */
  uint64_t seed;                      /* noise of inserted primitives */
} dp_idbastar_options;

typedef struct dp_idbastar_info {
  int32_t solved, exit_criteria;      /* DP_IDB_* */
  int32_t iterations, solutions;
  double cost;                        /* best solution cost (gScore), 1e8 if none */
  double last_delta;
  int64_t last_primitives, primitives_added, expands;
  double time_ms;
} dp_idbastar_info;

void dp_idbastar_options_default(dp_idbastar_options* opt);
/* `library`: the primitive set (host arrays, as dp_upload_primitives);
 * `search`: options of every dp_search run (delta is overridden).  The best
 * trajectory is available through dp_search_trajectory afterwards. */
int dp_idbastar(dp_ctx* ctx, const dp_prims* library, const double* start, const double* goal,
                const dp_search_options* search, const dp_idbastar_options* opt, dp_idbastar_info* info);

/* ---- the other callers of the expansion kernels -----------------------------
 * tdbastar (src/tdbastar/tdbastar.cpp:313-447, 778-905) and dbrrt
 * (src/dbrrt/dbrrt.cpp:580-660) run the same per-pair work as dbastar: the
 * applicable primitives of a node, placed at it, checked for bounds and
 * collisions.  They differ in the order of the applicable list, the search
 * direction, tdbastar's constraints and dbrrt's choice of one primitive per
 * node.  The first three are settings of a context that every later
 * dp_expand follows; dbrrt's choice is dp_expand_select. */

/* Pair order inside each node.  Expander::expand_lazy shuffles the
 * applicable list (dbastar.hpp:302-303, tdbastar.hpp:276-277) and dbrrt
 * reshuffles a node's stored list (dbrrt.cpp:611).  order is [rows][P] (P
 * the uploaded primitive count), each row a permutation of 0..P-1: node i's
 * applicable primitives come in the order they appear in row (i mod rows),
 * and that order is the batch's pair order everywhere (duplicates, branching
 * cap, dp_fetch_pairs).  rows = 0 (order NULL) restores ascending order;
 * dp_upload_primitives also restores it.  DP_ERR_ARG if a row is not a
 * permutation. */
int dp_set_order(dp_ctx* ctx, const int32_t* order, int32_t rows);

/* Search direction, tdbastar's reverse_search (tdbastar.cpp:575-576, 792).
 * forward = 0: applicability is keyed on each primitive's last state
 * (nigh_factory_t, nigh_custom_spaces.hpp:576-586), the primitive is placed
 * with offset - last position so that it ends at the node
 * (tdbastar.hpp:193-196), and the new node's state is the placed
 * primitive's first state (check_goal, tdbastar.cpp:453-458).  Planar
 * systems only (the reference assumes a 2-D offset); DP_ERR_ARG for the
 * quadrotor.  forward = 1 restores dbastar's direction. */
int dp_set_direction(dp_ctx* ctx, int32_t forward);

/* tdbastar's constraints (Constraint, tdbastar.hpp:155-158): n states
 * [n][nx] at times [n].  Every pair that passes the bounds and collision
 * checks is tested against them (tdbastar.cpp:402-443): reaches_goal =
 * distance(first placed state, or the last one backward, goal) <= delta;
 * per constraint, in list order, the time index
 * lround((float)(time - (float)g_parent) / ref_dt) selects state k when
 * 0 <= k < L-1, or the last state when k >= L-1 and reaches_goal; the pair is
 * violating when distance(state, constrained state) <= delta (delta and goal
 * from dp_expand_params).  A violating pair gets DP_V_CONSTRAINT, first_k =
 * the state index of the first violated constraint, and no duplicate test or
 * cap count.  n = 0 clears.  Planar systems only (DP_ERR_ARG for the
 * quadrotor). */
int dp_set_constraints(dp_ctx* ctx, const double* times, const double* states, int32_t n);

/* dbrrt's expansion step (dbrrt.cpp:580-660) for each uploaded node: its
 * applicable primitives (nearestR(canonical node, radius) over first states;
 * the Expander is built with options_dbrrt.delta, dbrrt.cpp:413-414), in
 * the context's pair order, each checked by check_lazy_trajectory (bounds of
 * every state, then collision; dbastar.cpp:391-490); of the valid ones the
 * first with the smallest distance(last state, target) (strict <,
 * dbrrt.cpp:627-642), or the first valid one with choose_first.  Backward
 * expansion in dbrrt is this step over the reversed primitive set
 * (dp_reverse_primitives) in a second context. */
typedef struct dp_select_params {
  double radius;            /* applicability radius, options_dbrrt.delta */
  double cost_jump;         /* options_dbrrt.cost_jump */
  int32_t choose_first;     /* options_dbrrt.choose_first_motion_valid */
  int32_t reserved;
  double goal[DP_MAX_NX];   /* for h = lower_bound_time(new state, goal) */
} dp_select_params;

/* Per node, caller-owned host buffers (any pointer may be NULL). */
typedef struct dp_select_out {
  int32_t* prim;            /* chosen primitive, -1 when no applicable primitive is valid */
  int64_t* pair;            /* its index in the batch's pair order (dp_fetch_pairs), -1 when none */
  double* distance;         /* distance(last placed state, target), NaN when none */
  double* state;            /* [N][nx] last placed state (__expand_end), NaN when none */
  double* start;            /* [N][nx] first placed state (__expand_start) */
  double* g;                /* g_parent + motion cost + cost_jump * lower_bound_time(node, start) */
  double* h;                /* lower_bound_time(state, goal) */
} dp_select_out;

/* targets: host [N][nx], one per uploaded node.  Enqueued on `stream` (NULL
 * = the context stream); dp_fetch_select waits for it.  Verdicts of every
 * pair are available through dp_fetch_pairs (resolution DP_R_NONE: no
 * duplicate test runs) and dp_get_counts. */
int dp_expand_select(dp_ctx* ctx, const double* targets, const dp_select_params* params, void* stream);
int dp_fetch_select(dp_ctx* ctx, dp_select_out* out);

/* reverse_motions (dbrrt.cpp:252-339) for the unicycles: each primitive
 * translated by minus its last position, states and actions reversed, the
 * first and last states perturbed by noise * U[-1, 1] drawn from a seeded
 * stream (the reference draws 1e-7 * Eigen::Random; noise = 0 disables it),
 * cost = ref_dt * #actions.  DP_ERR_ARG for other systems (the reference
 * supports only translation-invariant ones).  Release with
 * dp_free_primitives. */
int dp_reverse_primitives(const dp_prims* in, int32_t system, double noise, uint64_t seed, dp_prim_set* out);

/* ---- many planning instances in one launch ---------------------------------
 * No counterpart in the reference, which plans one instance per process.  A
 * batch groups n contexts (independent instances, any mix of systems, one
 * device); dp_batch_expand expands the uploaded nodes of all of them.
 * Each instance's pipeline up to its conflict lists runs on its own stream;
 * the ordered walks of all instances then run as one work queue per system
 * (waves take (instance, node) items interleaved over the instances).
 * DP_BATCH_GRAPH: all of it as one hipGraph launch, captured on the first
 * call and again whenever a context's node count, primitives, world, visited
 * set, buffer capacities or surrogate, or the params, changed since
 * (replayed otherwise; a capture call also runs the batch once eagerly).
 * DP_BATCH_STREAMS: the same work launched directly.  Results are read per
 * context (dp_get_counts, dp_fetch_pairs); per-stage dp_last_timing is not
 * kept for batch launches. */
enum { DP_BATCH_GRAPH = 0, DP_BATCH_STREAMS = 1 };
typedef struct dp_batch dp_batch;
dp_batch* dp_batch_create(dp_ctx* const* ctxs, int32_t n, int32_t mode);
void dp_batch_destroy(dp_batch* b);
/* params[n], one per context; stream NULL = the batch's own stream. */
int dp_batch_expand(dp_batch* b, const dp_expand_params* params, void* stream);
/* Completion time of each instance of the last dp_batch_expand (its walk's
 * last node), ms[n] after the batch started on the GPU (constant-rate GPU
 * clock).  Waits for it. */
int dp_batch_latency(dp_batch* b, float* ms);
/* Graph captures and launches so far. */
int dp_batch_stats(const dp_batch* b, int64_t* captures, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif /* DYNOPLAN_AMD_H */
