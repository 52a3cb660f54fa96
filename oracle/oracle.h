/*
 * oracle.h -- CPU FP64 restatement of Dynoplan's db-A* node-expansion path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (dynoplan_amd/) links,
 * imports or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker and as the timed CPU
 * baseline.  The reference itself (C++, fcl 0.7, OMPL, Eigen, Boost,
 * dynobench, nigh) cannot be built in this image, so this restatement is
 * the oracle; every function cites the reference lines it follows.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - unicycle1 dynamics + primitive format: reference data/ fixtures
 *     (msgpack <-> boost .bin twins), committed golden vectors.
 *   - quadrotor dynamics structure: benchmark_initguess trajectories.
 *   - collision: closed forms, brute force, invariants (translation,
 *     permutation, monotonicity), a dense point-sampling cross-check.
 *   - metric / NN: brute force is exact by construction; grid == brute.
 *   - acos / sin-cos series: against libm to 2 ulp over the used range.
 *   - check_goal index rule: the reference's float expression, enumerated.
 * Parity unpinned (the values live in the absent dynobench/nigh): distance
 * weights, lower_bound_time speeds, state bounds, geometry sizes, nigh's
 * radius inclusivity (<= is used), fcl's contact-at-zero convention
 * (touching counts as collision), quadrotor constants and op order.
 */
#ifndef DP_ORACLE_H
#define DP_ORACLE_H

#include <stdint.h>
#include "../include/dynoplan_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_NN_BRUTE = 0, ORACLE_NN_GRID = 1 };
enum { ORACLE_COL_BRUTE = 0, ORACLE_COL_SWEPT = 1 };

/* Whole batch, canonical order (see include/dynoplan_amd.h).  visited is
 * row-major [nv][nx].  Outputs: out arrays sized >= number of pairs; returns
 * the number of pairs, or a negative DP_ERR_* code. */
int64_t oracle_expand(const dp_system_desc* sys, const dp_prims* prims, int32_t max_motions,
                      const dp_world* world, const double* visited, int64_t nv,
                      const double* nodes, const double* g_parent, int32_t N,
                      const dp_expand_params* par, int32_t nn_mode, int32_t col_mode,
                      dp_pairs_out* out, dp_counts* counts);

/* The same batch with the per-pair work spread over `nthreads` OpenMP
 * threads (identical results; used to time the multi-core CPU baseline). */
int64_t oracle_expand_mt(const dp_system_desc* sys, const dp_prims* prims, int32_t max_motions,
                         const dp_world* world, const double* visited, int64_t nv,
                         const double* nodes, const double* g_parent, int32_t N,
                         const dp_expand_params* par, int32_t nn_mode, int32_t col_mode,
                         dp_pairs_out* out, dp_counts* counts, int32_t nthreads);

/* Settings of the other callers of the expansion (NULL = dbastar's):
 * order: [order_rows][P], node i's pairs in the order of row i % order_rows
 * (the shuffled applicable list, dbastar.hpp:302-303); backward: tdbastar's
 * reverse search (applicability on last states, primitive placed to end at
 * the node, new state = first placed state); constraints: tdbastar's
 * (time, state) list checked on every valid pair (DP_V_CONSTRAINT). */
typedef struct oracle_ext {
  const int32_t* order;
  int32_t order_rows;
  int32_t backward;
  int32_t n_constraints;
  int32_t reserved;
  const double* con_times;   /* [n_constraints] */
  const double* con_states;  /* [n_constraints][nx] */
} oracle_ext;

int64_t oracle_expand_ext(const dp_system_desc* sys, const dp_prims* prims, int32_t max_motions,
                          const dp_world* world, const double* visited, int64_t nv,
                          const double* nodes, const double* g_parent, int32_t N,
                          const dp_expand_params* par, int32_t nn_mode, int32_t col_mode,
                          const oracle_ext* ext, dp_pairs_out* out, dp_counts* counts);

/* tdbastar's constraint time index (tdbastar.cpp:417-427). */
int32_t oracle_constraint_index(double time, double g, double dt, int32_t L1, int32_t reaches_goal);

/* dbrrt's expansion step per node (dbrrt.cpp:580-660); see oracle.c. */
int64_t oracle_select(const dp_system_desc* sys, const dp_prims* prims, int32_t max_motions, const dp_world* world,
                      const double* nodes, const double* g_parent, int32_t N, double radius, const double* targets,
                      int32_t choose_first, double cost_jump, const double* goal, const int32_t* order,
                      int32_t order_rows, int32_t* out_prim, int64_t* out_pair, double* out_dist, double* out_state,
                      double* out_start, double* out_g, double* out_h);

/* Radius query over the visited set (brute force, ascending CSR). */
int64_t oracle_radius(const dp_system_desc* sys, const double* visited, int64_t nv, const double* queries,
                      int64_t n, double radius, int64_t* out_offsets, int64_t* out_idx, int64_t idx_cap);

/* Building blocks, exported for the pinning tests. */
double oracle_distance(const dp_system_desc* sys, const double* a, const double* b);
double oracle_lower_bound_time(const dp_system_desc* sys, const double* a, const double* b);
int oracle_state_valid(const dp_system_desc* sys, const dp_world* world, const double* s);
int oracle_state_collides(const dp_system_desc* sys, const dp_world* world, const double* s);
int oracle_obb_aabb(double cx, double cy, double c, double s, double hl, double hw,
                    double ox, double oy, double hx, double hy);
void oracle_quad_step(const dp_system_desc* sys, const double* s, const double* u, double* out);
void oracle_unicycle1_step(const dp_system_desc* sys, const double* s, const double* u, double* out);
double oracle_acos01(double x);
void oracle_cos_sinc(double z, double* c, double* sinc);
int32_t oracle_goal_index(int32_t nn, int32_t num_check_goal, int32_t size);

/* Rollout of primitive `j` translated to `node` (transform_primitive);
 * writes (L+1) states row-major into out.  Returns L+1. */
int32_t oracle_transform_primitive(const dp_system_desc* sys, const dp_prims* prims, int32_t j,
                                   const double* node, double* out);

#ifdef __cplusplus
}
#endif
#endif
