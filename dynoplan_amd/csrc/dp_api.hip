// dp_api.hip -- the C-ABI of include/dynoplan_amd.h.
//
// Host side of the expansion path: owns device memory, converts the
// row-major inputs of the reference's data structures (Motion::traj,
// dynobench env, AStarNode::state_eig) into the SoA layouts the kernels
// stream, and enqueues the kernel pipeline of k_expand.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dp_internal.h"

static thread_local std::string g_err;

void dp_set_error(const std::string& msg) { g_err = msg; }

#define HIP_CHECK(expr)                                                         \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      dp_set_error(std::string(#expr) + ": " + hipGetErrorString(e_));          \
      return DP_ERR_HIP;                                                        \
    }                                                                           \
  } while (0)

#define ARG_CHECK(cond, msg)   \
  do {                         \
    if (!(cond)) {             \
      dp_set_error(msg);       \
      return DP_ERR_ARG;       \
    }                          \
  } while (0)

template <typename T>
static int dev_alloc(T** p, int64_t n) {
  dp_bump_epoch();
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) {
    dp_set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    *p = nullptr;
    return DP_ERR_NOMEM;
  }
  return DP_OK;
}

template <typename T>
static void dev_free(T** p) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
}

// Series coefficients of the acos / cos / sinc evaluations (same recurrences
// as written in dp_device.h's users).
static void fill_series(dp_dev_sys& h) {
  double a = 1.0;
  h.asin_c[0] = 1.0;
  for (int n = 1; n < DP_ASIN_TERMS; ++n) {
    a = a * (double)(2 * n - 1) / (double)(2 * n);
    h.asin_c[n] = a / (double)(2 * n + 1);
  }
  h.cos_c[0] = 1.0;
  h.sinc_c[0] = 1.0;
  for (int n = 1; n < DP_TRIG_TERMS; ++n) {
    h.cos_c[n] = -h.cos_c[n - 1] / (double)((2 * n - 1) * (2 * n));
    h.sinc_c[n] = -h.sinc_c[n - 1] / (double)((2 * n) * (2 * n + 1));
  }
}

static void build_dev_sys(dp_ctx* c) {
  dp_dev_sys& h = c->hsys;
  const dp_system_desc& d = c->desc;
  h.system = d.system;
  h.nx = d.nx;
  h.nu = d.nu;
  h.pos_dim = d.pos_dim;
  h.dt = d.dt;
  for (int k = 0; k < DP_MAX_NX; ++k) {
    h.lo[k] = d.x_lb[k];
    h.hi[k] = d.x_ub[k];
  }
  for (int k = 0; k < d.pos_dim; ++k) {
    h.lo[k] = c->world_lo[k];
    h.hi[k] = c->world_hi[k];
  }
  for (int k = 0; k < 4; ++k) h.w[k] = d.dist_w[k];
  h.max_vel = d.max_vel;
  h.max_w = d.max_angular_vel;
  h.max_acc = d.max_acc;
  h.max_wacc = d.max_angular_acc;
  h.body_hl = d.body_half[0];
  h.body_hw = d.body_half[1];
  h.trailer_hl = d.trailer_half[0];
  h.trailer_hw = d.trailer_half[1];
  h.hitch = d.hitch;
  h.radius = d.body_half[0];
  for (int k = 0; k < 8; ++k) h.quad[k] = d.quad[k];
  h.near_eps = (float)d.near_eps;
  fill_series(h);
}

static int upload_sys(dp_ctx* c) {
  build_dev_sys(c);
  HIP_CHECK(hipMemcpy(c->d_sys, &c->hsys, sizeof(dp_dev_sys), hipMemcpyHostToDevice));
  return DP_OK;
}

extern "C" {

const char* dp_last_error(void) { return g_err.c_str(); }

int dp_system_default(int32_t system, dp_system_desc* o) {
  ARG_CHECK(o, "null output");
  std::memset(o, 0, sizeof(*o));
  const double inf = INFINITY;
  for (int k = 0; k < DP_MAX_NX; ++k) {
    o->x_lb[k] = -inf;
    o->x_ub[k] = inf;
  }
  o->system = system;
  o->near_eps = 1e-4;
  switch (system) {
    case DP_SYS_UNICYCLE1:
      o->nx = 3; o->nu = 2; o->pos_dim = 2; o->dt = 0.1;
      o->dist_w[0] = 1.0; o->dist_w[1] = 0.5;
      o->max_vel = 0.5; o->max_angular_vel = 0.5;
      o->body_half[0] = 0.25; o->body_half[1] = 0.125;
      return DP_OK;
    case DP_SYS_UNICYCLE2:
      o->nx = 5; o->nu = 2; o->pos_dim = 2; o->dt = 0.1;
      o->x_lb[3] = -0.5; o->x_ub[3] = 0.5; o->x_lb[4] = -0.5; o->x_ub[4] = 0.5;
      o->dist_w[0] = 1.0; o->dist_w[1] = 0.5; o->dist_w[2] = 0.25; o->dist_w[3] = 0.25;
      o->max_vel = 0.5; o->max_angular_vel = 0.5; o->max_acc = 0.25; o->max_angular_acc = 0.25;
      o->body_half[0] = 0.25; o->body_half[1] = 0.125;
      return DP_OK;
    case DP_SYS_CAR_TRAILER:
      o->nx = 4; o->nu = 2; o->pos_dim = 2; o->dt = 0.1;
      o->dist_w[0] = 1.0; o->dist_w[1] = 0.5; o->dist_w[2] = 0.5;
      o->max_vel = 0.5; o->max_angular_vel = 0.5;
      o->body_half[0] = 0.25; o->body_half[1] = 0.125;
      o->trailer_half[0] = 0.15; o->trailer_half[1] = 0.125;
      o->hitch = 0.5; o->wheelbase = 0.25;
      return DP_OK;
    case DP_SYS_QUADROTOR:
      o->nx = 13; o->nu = 4; o->pos_dim = 3; o->dt = 0.01;
      for (int k = 7; k < 10; ++k) { o->x_lb[k] = -4.0; o->x_ub[k] = 4.0; }
      for (int k = 10; k < 13; ++k) { o->x_lb[k] = -8.0; o->x_ub[k] = 8.0; }
      o->dist_w[0] = 1.0; o->dist_w[1] = 0.5; o->dist_w[2] = 0.1; o->dist_w[3] = 0.05;
      o->max_vel = 4.0; o->max_acc = 8.0; o->max_angular_vel = 8.0;
      o->body_half[0] = 0.1;
      o->quad[0] = 0.034;                   /* mass (kg) */
      o->quad[1] = 9.81;                    /* g */
      o->quad[2] = 0.046 * 0.7071067811865476; /* X-config moment arm (m) */
      o->quad[3] = 0.006;                   /* torque / thrust ratio */
      o->quad[4] = 16.571710e-6;            /* Jx */
      o->quad[5] = 16.655602e-6;            /* Jy */
      o->quad[6] = 29.261652e-6;            /* Jz */
      return DP_OK;
  }
  dp_set_error("unknown system id");
  return DP_ERR_ARG;
}

dp_ctx* dp_create(const dp_system_desc* sys, int32_t device) {
  if (!sys) {
    dp_set_error("null system description");
    return nullptr;
  }
  if (sys->nx <= 0 || sys->nx > DP_MAX_NX || (sys->pos_dim != 2 && sys->pos_dim != 3) ||
      !(sys->dist_w[0] > 0.0) || !(sys->dt > 0.0)) {
    dp_set_error("invalid system description");
    return nullptr;
  }
  static const int kNx[4] = {3, 5, 13, 4};
  if (sys->system < 0 || sys->system > 3 || sys->nx != kNx[sys->system]) {
    dp_set_error("system id / nx mismatch");
    return nullptr;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    dp_set_error("no HIP device " + std::to_string(device));
    return nullptr;
  }
  if (hipSetDevice(device) != hipSuccess) {
    dp_set_error("hipSetDevice failed");
    return nullptr;
  }
  dp_ctx* c = new dp_ctx();
  c->device = device;
  c->desc = *sys;
  c->nbox = sys->system == DP_SYS_CAR_TRAILER ? 2 : 1;
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
  for (int k = 0; k < DP_EVENTS && ok; ++k) ok = hipEventCreate(&c->ev[k]) == hipSuccess;
  ok = ok && dev_alloc(&c->d_sys, 1) == DP_OK;
  ok = ok && dev_alloc(&c->d_counters, DPC_COUNT) == DP_OK;
  ok = ok && dev_alloc(&c->d_sticky, DPS_COUNT) == DP_OK;
  ok = ok && hipMemset(c->d_sticky, 0, sizeof(long long) * DPS_COUNT) == hipSuccess;
  ok = ok && dev_alloc(&c->d_wshift, 2) == DP_OK;
  ok = ok && hipMemset(c->d_wshift, 0, sizeof(int32_t) * 2) == hipSuccess;
  ok = ok && dev_alloc(&c->d_lneed, 1) == DP_OK;
  ok = ok && hipMemset(c->d_lneed, 0, sizeof(int32_t)) == hipSuccess;
  ok = ok && dev_alloc(&c->d_total, 1) == DP_OK;
  ok = ok && hipMemset(c->d_counters, 0, sizeof(long long) * DPC_COUNT) == hipSuccess;
  ok = ok && hipMemset(c->d_total, 0, sizeof(int32_t)) == hipSuccess;
  ok = ok && upload_sys(c) == DP_OK;
  ok = ok && hipHostMalloc((void**)&c->h_pinned, sizeof(long long) * 4) == hipSuccess;
  ok = ok && dp_noalloc_init_device(device) == DP_OK;
  if (!ok) {
    if (g_err.empty()) dp_set_error("context creation failed");
    dp_destroy(c);
    return nullptr;
  }
  return c;
}

void dp_destroy(dp_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dp_mlp_free(c);
  dev_free(&c->d_sys);
  dev_free(&c->d_wshift);
  dev_free(&c->d_lneed);
  dev_free(&c->d_trace);
  dev_free(&c->d_node_obs);
  dev_free(&c->d_node_nobs);
  dev_free(&c->d_off);
  dev_free(&c->d_st);
  dev_free(&c->d_start);
  dev_free(&c->d_last);
  dev_free(&c->d_start32);
  dev_free(&c->d_last32);
  dev_free(&c->d_cost);
  dev_free(&c->d_order);
  dev_free(&c->d_con);
  dev_free(&c->d_targets);
  dev_free(&c->d_sel_q);
  dev_free(&c->d_sel);
  dev_free(&c->d_geo);
  dev_free(&c->d_sweep);
  dev_free(&c->d_srange);
  dev_free(&c->d_act);
  dev_free(&c->d_act32);
  dev_free(&c->d_goalidx);
  dev_free(&c->d_obs);
  dev_free(&c->d_vis);
  dev_free(&c->d_vis_sorted);
  dev_free(&c->d_vis_cell);
  dev_free(&c->d_vis_perm);
  dp_radius_free(c);
  dev_free(&c->d_vis_start);
  dev_free(&c->d_nodes);
  dev_free(&c->d_gpar);
  dev_free(&c->d_app);
  dev_free(&c->d_app_mask);
  dev_free(&c->d_node_novel);
  dev_free(&c->d_pnode);
  dev_free(&c->d_pprim);
  dev_free(&c->d_verdict);
  dev_free(&c->d_firstk);
  dev_free(&c->d_goal);
  dev_free(&c->d_res);
  dev_free(&c->d_flags);
  dev_free(&c->d_final);
  dev_free(&c->d_g);
  dev_free(&c->d_h);
  dev_free(&c->d_cont_start);
  dev_free(&c->d_cont_cursor);
  dev_free(&c->d_cont_idx);
  dev_free(&c->d_cont_x);
  dev_free(&c->d_conf_cnt);
  dev_free(&c->d_conf);
  dev_free(&c->d_conf_ov);
  dev_free(&c->d_conf_pool);
  dev_free(&c->d_tiles);
  dev_free(&c->d_node_cursor);
  dev_free(&c->d_node_count);
  dev_free(&c->d_chunk);
  dev_free(&c->d_q64);
  dev_free(&c->d_qperm);
  dev_free(&c->d_chunk_node);
  dev_free(&c->d_nshort);
  dev_free(&c->d_rot64);
  dev_free(&c->d_preach);
  dev_free(&c->d_dv64);
  dev_free(&c->d_dv32);
  dev_free(&c->d_rbad64);
  dev_free(&c->d_rbad32);
  dev_free(&c->d_rnear32);
  dev_free(&c->d_qrec);
  if (c->h_pinned) (void)hipHostFree(c->h_pinned);
  c->h_pinned = nullptr;
  dev_free(&c->d_counters);
  dev_free(&c->d_sticky);
  dev_free(&c->d_total);
  for (int k = 0; k < DP_EVENTS; ++k)
    if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

void* dp_stream(dp_ctx* c) { return c ? (void*)c->stream : nullptr; }

int dp_upload_primitives(dp_ctx* c, const dp_prims* pr, int32_t max_motions) {
  ARG_CHECK(c && pr && pr->offsets && pr->states, "null argument");
  ARG_CHECK(pr->nx == c->desc.nx, "primitive nx does not match the system");
  ARG_CHECK(pr->offsets[0] == 0, "offsets[0] must be 0");
  HIP_CHECK(hipSetDevice(c->device));
  int32_t P = pr->P;
  if (max_motions > 0 && max_motions < P) P = max_motions;  // dbastar.cpp:566-567
  ARG_CHECK(P > 0, "no primitives");
  const int nx = c->desc.nx;
  const bool quad = c->desc.system == DP_SYS_QUADROTOR;
  if (quad) ARG_CHECK(pr->actions && pr->nu == c->desc.nu, "quadrotor primitives need actions with nu=4");
  std::vector<int32_t> off(P + 1);
  int32_t maxlen = 0;
  for (int32_t j = 0; j <= P; ++j) off[j] = pr->offsets[j];
  for (int32_t j = 0; j < P; ++j) {
    int32_t l = off[j + 1] - off[j];
    ARG_CHECK(l >= 1, "primitive with no state");
    maxlen = std::max(maxlen, l);
  }
  const int32_t S = off[P];
  const double* st = pr->states;
  std::vector<double> soa((size_t)nx * S), start((size_t)nx * P), last((size_t)nx * P), cost((size_t)P);
  for (int32_t s = 0; s < S; ++s)
    for (int k = 0; k < nx; ++k) soa[(size_t)k * S + s] = st[(int64_t)s * nx + k];
  for (int32_t j = 0; j < P; ++j)
    for (int k = 0; k < nx; ++k) {
      start[(size_t)k * P + j] = st[(int64_t)off[j] * nx + k];
      last[(size_t)k * P + j] = st[(int64_t)(off[j + 1] - 1) * nx + k];
    }
  // Motion::cost (dbrrt.cpp:645-648): as given, else ref_dt per action
  for (int32_t j = 0; j < P; ++j) cost[j] = pr->cost ? pr->cost[j] : (off[j + 1] - off[j] - 1) * c->desc.dt;

  // Footprint geometry per state (compute_col_shape, robots.cpp:3190-3215):
  // translation leaves orientation unchanged, so cos/sin and the trailer's
  // offset from the state position are per-primitive constants.
  const int nbox = c->nbox;
  std::vector<double> geo;
  std::vector<double> sweep((size_t)4 * P);
  if (!quad) {
    geo.assign((size_t)4 * nbox * S, 0.0);
    const double hitch = c->desc.hitch;
    for (int32_t s = 0; s < S; ++s) {
      const double* x = st + (int64_t)s * nx;
      double c0 = std::cos(x[2]), s0 = std::sin(x[2]);
      geo[(size_t)2 * S + s] = c0;
      geo[(size_t)3 * S + s] = s0;
      if (nbox > 1) {
        double c1 = std::cos(x[3]), s1 = std::sin(x[3]);
        geo[(size_t)4 * S + s] = -(hitch * c1);
        geo[(size_t)5 * S + s] = -(hitch * s1);
        geo[(size_t)6 * S + s] = c1;
        geo[(size_t)7 * S + s] = s1;
      }
    }
    for (int32_t j = 0; j < P; ++j) {
      double lo0 = INFINITY, hi0 = -INFINITY, lo1 = INFINITY, hi1 = -INFINITY;
      for (int32_t s = off[j]; s < off[j + 1]; ++s) {
        for (int b = 0; b < nbox; ++b) {
          double cx = st[(int64_t)s * nx] + (b ? geo[(size_t)4 * S + s] : 0.0);
          double cy = st[(int64_t)s * nx + 1] + (b ? geo[(size_t)5 * S + s] : 0.0);
          double cc = std::fabs(geo[(size_t)(4 * b + 2) * S + s]), ss = std::fabs(geo[(size_t)(4 * b + 3) * S + s]);
          double hl = b ? c->desc.trailer_half[0] : c->desc.body_half[0];
          double hw = b ? c->desc.trailer_half[1] : c->desc.body_half[1];
          double ex = hl * cc + hw * ss, ey = hl * ss + hw * cc;
          lo0 = std::min(lo0, cx - ex);
          hi0 = std::max(hi0, cx + ex);
          lo1 = std::min(lo1, cy - ey);
          hi1 = std::max(hi1, cy + ey);
        }
      }
      sweep[j] = lo0;
      sweep[(size_t)P + j] = hi0;
      sweep[(size_t)2 * P + j] = lo1;
      sweep[(size_t)3 * P + j] = hi1;
    }
    c->reach_box[0] = c->reach_box[2] = INFINITY;
    c->reach_box[1] = c->reach_box[3] = -INFINITY;
    for (int32_t j = 0; j < P; ++j) {
      c->reach_box[0] = std::min(c->reach_box[0], sweep[j]);
      c->reach_box[1] = std::max(c->reach_box[1], sweep[(size_t)P + j]);
      c->reach_box[2] = std::min(c->reach_box[2], sweep[(size_t)2 * P + j]);
      c->reach_box[3] = std::max(c->reach_box[3], sweep[(size_t)3 * P + j]);
    }
    // The reverse search places primitive j at node - last_j: its swept box
    // relative to the node is sweep_j - last_j.  The pair test rounds
    // sweep_j + (node - last_j) instead, a few ulps away: the union box is
    // widened by 1e-6 (relative to the extents) to stay conservative.
    double rb[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
    double ext = 1.0;
    for (int32_t j = 0; j < P; ++j) {
      const double lx = last[j], ly = last[(size_t)P + j];
      rb[0] = std::min(rb[0], sweep[j] - lx);
      rb[1] = std::max(rb[1], sweep[(size_t)P + j] - lx);
      rb[2] = std::min(rb[2], sweep[(size_t)2 * P + j] - ly);
      rb[3] = std::max(rb[3], sweep[(size_t)3 * P + j] - ly);
      ext = std::max({ext, std::fabs(sweep[j]), std::fabs(sweep[(size_t)P + j]), std::fabs(sweep[(size_t)2 * P + j]),
                      std::fabs(sweep[(size_t)3 * P + j]), std::fabs(lx), std::fabs(ly)});
    }
    const double wpad = 1e-6 * ext;
    c->reach_box_bwd[0] = rb[0] - wpad;
    c->reach_box_bwd[1] = rb[1] + wpad;
    c->reach_box_bwd[2] = rb[2] - wpad;
    c->reach_box_bwd[3] = rb[3] + wpad;
  }
  // Per-primitive range of every state component: the check kernel proves
  // "no state out of bounds" from it without reading the states.  A NaN
  // component gives a NaN range, which fails every comparison and so never
  // proves anything.
  std::vector<double> srange((size_t)2 * nx * P);
  for (int32_t j = 0; j < P; ++j) {
    for (int k = 0; k < nx; ++k) {
      double mn = INFINITY, mx = -INFINITY;
      bool nan = false;
      for (int32_t s = off[j]; s < off[j + 1]; ++s) {
        const double v = st[(int64_t)s * nx + k];
        nan = nan || v != v;
        mn = std::min(mn, v);
        mx = std::max(mx, v);
      }
      if (nan) mn = mx = NAN;
      srange[(size_t)k * P + j] = mn;
      srange[((size_t)nx + k) * P + j] = mx;
    }
  }
  int rc;
  if ((rc = dev_alloc(&c->d_off, P + 1))) return rc;
  if ((rc = dev_alloc(&c->d_st, (int64_t)nx * S))) return rc;
  if ((rc = dev_alloc(&c->d_start, (int64_t)nx * P))) return rc;
  if ((rc = dev_alloc(&c->d_last, (int64_t)nx * P))) return rc;
  if ((rc = dev_alloc(&c->d_cost, P))) return rc;
  if ((rc = dev_alloc(&c->d_sweep, (int64_t)4 * P))) return rc;
  if ((rc = dev_alloc(&c->d_srange, (int64_t)2 * nx * P))) return rc;
  if ((rc = dev_alloc(&c->d_goalidx, (int64_t)8 * P))) return rc;
  HIP_CHECK(hipMemcpy(c->d_off, off.data(), sizeof(int32_t) * (P + 1), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->d_st, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->d_start, start.data(), sizeof(double) * start.size(), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->d_last, last.data(), sizeof(double) * last.size(), hipMemcpyHostToDevice));
  {
    std::vector<float> s32(start.begin(), start.end()), l32(last.begin(), last.end());
    if ((rc = dev_alloc(&c->d_start32, (int64_t)nx * P))) return rc;
    if ((rc = dev_alloc(&c->d_last32, (int64_t)nx * P))) return rc;
    HIP_CHECK(hipMemcpy(c->d_start32, s32.data(), sizeof(float) * s32.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->d_last32, l32.data(), sizeof(float) * l32.size(), hipMemcpyHostToDevice));
  }
  HIP_CHECK(hipMemcpy(c->d_cost, cost.data(), sizeof(double) * P, hipMemcpyHostToDevice));
  dev_free(&c->d_order);  // a pair order refers to the previous primitive count
  c->order_rows = 0;
  HIP_CHECK(hipMemcpy(c->d_sweep, sweep.data(), sizeof(double) * sweep.size(), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->d_srange, srange.data(), sizeof(double) * srange.size(), hipMemcpyHostToDevice));
  if (!quad) {
    if ((rc = dev_alloc(&c->d_geo, (int64_t)geo.size()))) return rc;
    HIP_CHECK(hipMemcpy(c->d_geo, geo.data(), sizeof(double) * geo.size(), hipMemcpyHostToDevice));
  } else {
    const int nu = c->desc.nu;
    const int64_t A = (int64_t)S - P;
    const double* ac = pr->actions;
    std::vector<double> asoa((size_t)nu * (A > 0 ? A : 1));
    std::vector<float> asoa32(asoa.size());
    for (int32_t j = 0; j < P; ++j) {
      const int64_t src0 = (int64_t)(pr->offsets[j] - j);  // first action row of primitive j
      const int64_t dst0 = (int64_t)(off[j] - j);
      const int32_t L = off[j + 1] - off[j] - 1;
      for (int32_t k = 0; k < L; ++k)
        for (int u = 0; u < nu; ++u) {
          double v = ac[(src0 + k) * nu + u];
          asoa[(size_t)(dst0 + k) * nu + u] = v;  // one row of 4 per step: a single vector load
          asoa32[(size_t)(dst0 + k) * nu + u] = (float)v;
        }
    }
    if ((rc = dev_alloc(&c->d_act, (int64_t)asoa.size()))) return rc;
    if ((rc = dev_alloc(&c->d_act32, (int64_t)asoa32.size()))) return rc;
    HIP_CHECK(hipMemcpy(c->d_act, asoa.data(), sizeof(double) * asoa.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->d_act32, asoa32.data(), sizeof(float) * asoa32.size(), hipMemcpyHostToDevice));
    double reach = 0.0;
    std::vector<float> preach((size_t)P);
    for (int32_t j = 0; j < P; ++j) {
      double rj = 0.0;
      for (int32_t s = off[j]; s < off[j + 1]; ++s)
        for (int k = 0; k < 3; ++k)
          rj = std::max(rj, std::fabs(st[(int64_t)s * nx + k] - st[(int64_t)off[j] * nx + k]));
      preach[j] = std::nextafter((float)rj, INFINITY);  // never below the FP64 value
      reach = std::max(reach, rj);
    }
    c->prim_reach = reach;
    if ((rc = dev_alloc(&c->d_preach, P))) return rc;
    HIP_CHECK(hipMemcpy(c->d_preach, preach.data(), sizeof(float) * P, hipMemcpyHostToDevice));
  }
  dp_bump_epoch();
  c->h_st.assign(st, st + (size_t)S * nx);
  c->h_nu = pr->actions && pr->nu > 0 ? pr->nu : 0;
  if (c->h_nu) {
    // action rows of primitive j start at offsets[j] - j
    c->h_ac.assign((size_t)(S - P) * c->h_nu, 0.0);
    for (int32_t j = 0; j < P; ++j)
      for (int32_t k = 0; k < off[j + 1] - off[j] - 1; ++k)
        for (int u = 0; u < c->h_nu; ++u)
          c->h_ac[(size_t)(off[j] - j + k) * c->h_nu + u] = pr->actions[(int64_t)(pr->offsets[j] - j + k) * pr->nu + u];
  } else {
    c->h_ac.clear();
  }
  c->P = P;
  c->S = S;
  c->maxlen = maxlen;
  c->h_off = off;
  c->goal_num = -1;
  if (quad) {
    // attitude / rate trajectories shared by every pair of a primitive
    if ((rc = dev_alloc(&c->d_rot64, (int64_t)7 * S))) return rc;
    if ((rc = dev_alloc(&c->d_dv64, (int64_t)4 * S))) return rc;
    if ((rc = dev_alloc(&c->d_dv32, (int64_t)4 * S))) return rc;
    if ((rc = dev_alloc(&c->d_rbad64, P))) return rc;
    if ((rc = dev_alloc(&c->d_rbad32, P))) return rc;
    if ((rc = dev_alloc(&c->d_rnear32, P))) return rc;
    if ((rc = dev_alloc(&c->d_qrec, (int64_t)16 * P))) return rc;
    if ((rc = dp_launch_quad_rot(c, c->stream))) return rc;
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  return DP_OK;
}

int dp_upload_world(dp_ctx* c, const dp_world* w) {
  ARG_CHECK(c && w, "null argument");
  ARG_CHECK(w->n_boxes >= 0 && w->n_spheres >= 0, "negative obstacle count");
  ARG_CHECK(w->n_boxes + w->n_spheres <= DP_MAX_OBS, "more obstacles than the LDS staging limit (1024)");
  ARG_CHECK(w->n_boxes == 0 || w->boxes, "null boxes");
  ARG_CHECK(w->n_spheres == 0 || w->spheres, "null spheres");
  const bool planar = c->desc.pos_dim == 2;
  ARG_CHECK(!planar || w->n_spheres == 0, "planar systems take box obstacles only");
  for (int k = 0; k < c->desc.pos_dim; ++k) ARG_CHECK(w->hi[k] > w->lo[k], "empty environment bounds");
  HIP_CHECK(hipSetDevice(c->device));
  std::vector<double> obs;
  if (planar) {
    obs.resize((size_t)4 * w->n_boxes);
    for (int o = 0; o < w->n_boxes; ++o) {
      obs[4 * o + 0] = w->boxes[6 * o + 0];
      obs[4 * o + 1] = w->boxes[6 * o + 1];
      obs[4 * o + 2] = w->boxes[6 * o + 3];
      obs[4 * o + 3] = w->boxes[6 * o + 4];
    }
  } else {
    obs.assign(w->boxes, w->boxes + 6 * (size_t)w->n_boxes);
    obs.insert(obs.end(), w->spheres, w->spheres + 4 * (size_t)w->n_spheres);
  }
  int rc;
  if ((rc = dev_alloc(&c->d_obs, (int64_t)obs.size()))) return rc;
  if (!obs.empty()) HIP_CHECK(hipMemcpy(c->d_obs, obs.data(), sizeof(double) * obs.size(), hipMemcpyHostToDevice));
  c->nb = w->n_boxes;
  c->ns = w->n_spheres;
  for (int k = 0; k < 3; ++k) {
    c->world_lo[k] = w->lo[k];
    c->world_hi[k] = w->hi[k];
  }
  c->have_world = true;
  c->grid_dirty = true;
  dp_bump_epoch();
  return upload_sys(c);
}

static int ensure_visited_cap(dp_ctx* c, int64_t need) {
  if (need <= c->nv_cap) return DP_OK;
  int64_t cap = std::max<int64_t>(need, std::max<int64_t>(1024, c->nv_cap * 2));
  const int nx = c->desc.nx;
  double* nv_ = nullptr;
  dp_bump_epoch();
  HIP_CHECK(hipMalloc(&nv_, sizeof(double) * nx * cap));
  if (c->d_vis && c->nv > 0) {
    HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < nx; ++k)
      HIP_CHECK(hipMemcpy(nv_ + (size_t)k * cap, c->d_vis + (size_t)k * c->nv_cap, sizeof(double) * c->nv,
                          hipMemcpyDeviceToDevice));
  }
  dev_free(&c->d_vis);
  c->d_vis = nv_;
  int rc;
  if ((rc = dev_alloc(&c->d_vis_sorted, (int64_t)nx * cap))) return rc;
  if ((rc = dev_alloc(&c->d_vis_cell, cap))) return rc;
  if ((rc = dev_alloc(&c->d_vis_perm, cap))) return rc;
  c->nv_cap = cap;
  c->grid_dirty = true;
  return DP_OK;
}

int dp_visited_reset(dp_ctx* c) {
  ARG_CHECK(c, "null context");
  c->nv = 0;
  c->grid_dirty = true;
  dp_bump_epoch();
  return DP_OK;
}

int64_t dp_visited_size(const dp_ctx* c) { return c ? c->nv : -1; }

int dp_visited_append(dp_ctx* c, const double* states, int64_t n) {
  ARG_CHECK(c && (n == 0 || states) && n >= 0, "invalid argument");
  if (n == 0) return DP_OK;
  HIP_CHECK(hipSetDevice(c->device));
  HIP_CHECK(hipDeviceSynchronize());
  int rc = ensure_visited_cap(c, c->nv + n);
  if (rc) return rc;
  const int nx = c->desc.nx;
  std::vector<double> col((size_t)n);
  for (int k = 0; k < nx; ++k) {
    for (int64_t i = 0; i < n; ++i) col[i] = states[i * nx + k];
    HIP_CHECK(hipMemcpy(c->d_vis + (size_t)k * c->nv_cap + c->nv, col.data(), sizeof(double) * n,
                        hipMemcpyHostToDevice));
  }
  c->nv += n;
  c->grid_dirty = true;
  dp_bump_epoch();
  return DP_OK;
}

int dp_reserve(dp_ctx* c, int32_t max_nodes, int64_t max_pairs) {
  ARG_CHECK(c && max_nodes >= 0 && max_pairs >= 0, "invalid argument");
  HIP_CHECK(hipSetDevice(c->device));
  const int nx = c->desc.nx;
  int rc;
  if (max_nodes > c->N_cap) {
    HIP_CHECK(hipStreamSynchronize(c->stream));
    if ((rc = dev_alloc(&c->d_nodes, (int64_t)nx * max_nodes))) return rc;
    if ((rc = dev_alloc(&c->d_gpar, max_nodes))) return rc;
    if ((rc = dev_alloc(&c->d_app, (int64_t)max_nodes + 1))) return rc;
    if ((rc = dev_alloc(&c->d_node_novel, (int64_t)max_nodes + 1))) return rc;
    if ((rc = dev_alloc(&c->d_node_cursor, max_nodes))) return rc;
    if ((rc = dev_alloc(&c->d_node_count, max_nodes))) return rc;
    if ((rc = dev_alloc(&c->d_chunk, (int64_t)max_nodes + 1))) return rc;
    if ((rc = dev_alloc(&c->d_nshort, max_nodes))) return rc;
    if ((rc = dev_alloc(&c->d_node_obs, (int64_t)DP_NODE_OBS * max_nodes))) return rc;
    if ((rc = dev_alloc(&c->d_node_nobs, max_nodes))) return rc;
    c->N_cap = max_nodes;
  }
  if (max_pairs > c->pair_cap) {
    ARG_CHECK(max_pairs < (int64_t)INT32_MAX, "more than 2^31 pairs in one batch");
    HIP_CHECK(hipStreamSynchronize(c->stream));
    if ((rc = dev_alloc(&c->d_pnode, max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_pprim, max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_verdict, max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_firstk, max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_goal, max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_res, max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_flags, max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_final, (int64_t)nx * max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_g, max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_h, max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_cont_idx, max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_cont_x, (int64_t)nx * max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_conf_cnt, max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_conf, (int64_t)DP_CONF_SLOTS * max_pairs))) return rc;
    if ((rc = dev_alloc(&c->d_conf_ov, max_pairs))) return rc;
    c->conf_pool_cap = 2 * max_pairs;  // about 1 in 10 lists overflows, by ~15 entries, on dense batches
    if ((rc = dev_alloc(&c->d_conf_pool, c->conf_pool_cap))) return rc;
    if (c->desc.system == DP_SYS_QUADROTOR && (rc = dev_alloc(&c->d_q64, max_pairs))) return rc;
    if (c->desc.system == DP_SYS_QUADROTOR && (rc = dev_alloc(&c->d_qperm, 2 * max_pairs))) return rc;
    c->pair_cap = max_pairs;
  }
  return DP_OK;
}

int dp_upload_nodes(dp_ctx* c, const double* nodes, const double* g_parent, int32_t N) {
  ARG_CHECK(c && nodes && N > 0, "invalid argument");
  int rc = dp_reserve(c, N, 0);
  if (rc) return rc;
  const int nx = c->desc.nx;
  std::vector<double> soa((size_t)nx * c->N_cap, 0.0), g((size_t)N, 0.0);
  for (int32_t i = 0; i < N; ++i)
    for (int k = 0; k < nx; ++k) soa[(size_t)k * c->N_cap + i] = nodes[(int64_t)i * nx + k];
  if (g_parent)
    for (int32_t i = 0; i < N; ++i) g[i] = g_parent[i];
  HIP_CHECK(hipDeviceSynchronize());  // kernels of a previous batch may run on a caller stream
  HIP_CHECK(hipMemcpy(c->d_nodes, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->d_gpar, g.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  if (c->N != N) dp_bump_epoch();
  c->N = N;
  c->sel_valid = false;
  return DP_OK;
}

}  // extern "C"

// check_goal state indices per primitive (dbastar.cpp:499-500, float arithmetic)
int dp_prepare_goal_index(dp_ctx* c, const dp_expand_params* p, hipStream_t s) {
  if (c->goal_num == p->num_check_goal) return DP_OK;
  std::vector<int32_t> gi((size_t)8 * c->P, 0);
  for (int32_t j = 0; j < c->P; ++j) {
    const int32_t size = c->h_off[j + 1] - c->h_off[j];
    for (int nn = 0; nn < p->num_check_goal; ++nn) {
      float r = (float)(nn + 1) / (float)(p->num_check_goal + 1);
      gi[(size_t)nn * c->P + j] = (int32_t)(r * (float)size);
    }
  }
  HIP_CHECK(hipStreamSynchronize(s));
  HIP_CHECK(hipMemcpy(c->d_goalidx, gi.data(), sizeof(int32_t) * gi.size(), hipMemcpyHostToDevice));
  c->goal_num = p->num_check_goal;
  return DP_OK;
}

// Pair buffers.  When every (node, primitive) pair fits the reserved
// capacity nothing is synchronised; otherwise the applicable pairs are counted
// exactly first (one small device-to-host read) and the buffers grow to that
// count plus 1/8 headroom.  Small batches reserve N * P outright.
int dp_size_pairs(dp_ctx* c, const dp_expand_params* p, hipStream_t s) {
  const int64_t worst = (int64_t)c->N * c->P;
  const int64_t kOutright = 1ll << 22;
  int rc = dp_reserve(c, c->N, std::min(worst, kOutright));
  if (rc) return rc;
  if ((rc = dp_ensure_u64(&c->d_app_mask, &c->app_mask_cap, (int64_t)c->N * ((c->P + 63) / 64)))) return rc;
  if (c->desc.system == DP_SYS_QUADROTOR &&
      (rc = dp_ensure_i32(&c->d_chunk_node, &c->chunk_node_cap, (int64_t)c->N * ((c->P + 63) / 64 + 1))))
    return rc;
  if (worst <= c->pair_cap) return DP_OK;
  if ((rc = dp_launch_count(c, p, s))) return rc;
  long long total = 0;
  HIP_CHECK(hipMemcpyAsync(c->h_pinned, c->d_counters + DPC_PAIRS, sizeof(long long), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  std::memcpy(&total, c->h_pinned, sizeof(long long));
  if (total > c->pair_cap) {
    ARG_CHECK(total < (long long)INT32_MAX / 2, "more than 2^30 applicable pairs in one batch");
    rc = dp_reserve(c, c->N, std::min<int64_t>(worst, total + total / 8 + 1024));
    if (rc) return rc;
  }
  return DP_OK;
}

extern "C" {

int dp_expand(dp_ctx* c, const dp_expand_params* p, void* stream) {
  ARG_CHECK(c && p, "null argument");
  ARG_CHECK(p->alpha > 0.0f && p->alpha < 1.0f, "Alpha needs to be between 0 and 1!");  // dbastar.cpp:588-590
  ARG_CHECK(p->delta >= 0.0f, "delta < 0 is not implemented");                          // dbastar.cpp:592-594
  ARG_CHECK(p->num_check_goal >= 0 && p->num_check_goal <= 8, "num_check_goal must be in [0, 8]");
  ARG_CHECK(!(p->use_surrogate && !c->forward), "the surrogate scores forward placements only");
  if (c->P <= 0 || !c->have_world || c->N <= 0) {
    dp_set_error("upload primitives, world and nodes before dp_expand");
    return DP_ERR_STATE;
  }
  HIP_CHECK(hipSetDevice(c->device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  int rc = dp_prepare_goal_index(c, p, s);
  if (rc) return rc;
  rc = dp_size_pairs(c, p, s);
  if (rc) return rc;
  rc = dp_launch_expand(c, p, s);
  c->last_pairs = -1;
  c->sel_valid = false;
  if (!rc) ++c->n_batches;
  return rc;
}

int dp_set_order(dp_ctx* c, const int32_t* order, int32_t rows) {
  ARG_CHECK(c && rows >= 0 && (rows == 0 || order), "invalid argument");
  if (c->P <= 0) {
    dp_set_error("upload primitives before dp_set_order");
    return DP_ERR_STATE;
  }
  HIP_CHECK(hipSetDevice(c->device));
  if (rows == 0) {
    HIP_CHECK(hipDeviceSynchronize());
    dev_free(&c->d_order);
    c->order_rows = 0;
    dp_bump_epoch();
    return DP_OK;
  }
  const int32_t P = c->P;
  std::vector<uint8_t> seen((size_t)P);
  for (int32_t r = 0; r < rows; ++r) {
    std::fill(seen.begin(), seen.end(), 0);
    for (int32_t k = 0; k < P; ++k) {
      const int32_t j = order[(int64_t)r * P + k];
      ARG_CHECK(j >= 0 && j < P && !seen[j], "a row of the pair order is not a permutation of 0..P-1");
      seen[j] = 1;
    }
  }
  HIP_CHECK(hipDeviceSynchronize());  // a batch in flight may read the previous order
  int rc = dev_alloc(&c->d_order, (int64_t)rows * P);
  if (rc) return rc;
  HIP_CHECK(hipMemcpy(c->d_order, order, sizeof(int32_t) * (size_t)rows * P, hipMemcpyHostToDevice));
  c->order_rows = rows;
  return DP_OK;
}

int dp_set_direction(dp_ctx* c, int32_t forward) {
  ARG_CHECK(c, "null context");
  ARG_CHECK(forward || c->desc.system != DP_SYS_QUADROTOR, "the reverse search is defined for planar systems");
  c->forward = forward ? 1 : 0;
  dp_bump_epoch();
  return DP_OK;
}

int dp_set_constraints(dp_ctx* c, const double* times, const double* states, int32_t n) {
  ARG_CHECK(c && n >= 0 && (n == 0 || (times && states)), "invalid argument");
  ARG_CHECK(n == 0 || c->desc.system != DP_SYS_QUADROTOR, "constraints are defined for planar systems");
  HIP_CHECK(hipSetDevice(c->device));
  HIP_CHECK(hipDeviceSynchronize());  // a batch in flight may read the previous list
  const int nx = c->desc.nx;
  if (n > c->con_cap) {
    int rc = dev_alloc(&c->d_con, (int64_t)n * (nx + 1));
    if (rc) return rc;
    c->con_cap = n;
  }
  if (n > 0) {
    std::vector<double> h((size_t)n * (nx + 1));
    for (int32_t k = 0; k < n; ++k) h[k] = times[k];
    for (int64_t k = 0; k < (int64_t)n * nx; ++k) h[(size_t)n + k] = states[k];
    HIP_CHECK(hipMemcpy(c->d_con, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  }
  c->ncon = n;
  dp_bump_epoch();
  return DP_OK;
}

int dp_expand_select(dp_ctx* c, const double* targets, const dp_select_params* sp, void* stream) {
  ARG_CHECK(c && targets && sp, "null argument");
  ARG_CHECK(sp->radius >= 0.0, "negative radius");
  ARG_CHECK(c->desc.system != DP_SYS_QUADROTOR, "dp_expand_select is defined for planar systems");
  if (!c->forward || c->ncon > 0) {
    dp_set_error("dp_expand_select runs forward without constraints (dbrrt.cpp:580-660)");
    return DP_ERR_STATE;
  }
  if (c->P <= 0 || !c->have_world || c->N <= 0) {
    dp_set_error("upload primitives, world and nodes before dp_expand_select");
    return DP_ERR_STATE;
  }
  HIP_CHECK(hipSetDevice(c->device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  const int nx = c->desc.nx;
  int rc;
  if (c->targets_cap < c->N_cap) {
    if ((rc = dev_alloc(&c->d_targets, (int64_t)nx * c->N_cap))) return rc;
    c->targets_cap = c->N_cap;
  }
  if (c->sel_cap < c->N_cap) {
    if ((rc = dev_alloc(&c->d_sel_q, 2 * (int64_t)c->N_cap))) return rc;  // pair, then primitive
    if ((rc = dev_alloc(&c->d_sel, (int64_t)(2 * nx + 3) * c->N_cap))) return rc;
    c->sel_cap = c->N_cap;
  }
  std::vector<double> soa((size_t)nx * c->N_cap, 0.0);
  for (int32_t i = 0; i < c->N; ++i)
    for (int k = 0; k < nx; ++k) soa[(size_t)k * c->N_cap + i] = targets[(int64_t)i * nx + k];
  HIP_CHECK(hipStreamSynchronize(s));
  HIP_CHECK(hipMemcpy(c->d_targets, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
  dp_expand_params p{};
  p.delta = 1.0f;  // the duplicate radius is not used: no duplicate test runs
  p.alpha = 0.5f;
  p.limit_branching_factor = -1;
  for (int k = 0; k < DP_MAX_NX; ++k) p.goal[k] = sp->goal[k];
  c->select_mode = true;
  c->select_radius = sp->radius;
  c->select_cost_jump = sp->cost_jump;
  c->select_first = sp->choose_first ? 1 : 0;
  rc = dp_size_pairs(c, &p, s);
  if (!rc) rc = dp_launch_expand(c, &p, s);
  c->select_mode = false;
  c->last_pairs = -1;
  c->sel_valid = rc == DP_OK;
  if (!rc) ++c->n_batches;
  return rc;
}

int dp_fetch_select(dp_ctx* c, dp_select_out* o) {
  ARG_CHECK(c && o, "null argument");
  if (!c->sel_valid) {
    dp_set_error("no dp_expand_select since the last expansion");
    return DP_ERR_STATE;
  }
  dp_counts cnt;
  int rc = dp_get_counts(c, &cnt);
  if (rc) return rc;
  const int nx = c->desc.nx;
  const int32_t N = c->N;
  std::vector<int32_t> q((size_t)N);
  std::vector<double> v((size_t)N * (2 * nx + 3));
  HIP_CHECK(hipMemcpy(q.data(), c->d_sel_q, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(v.data(), c->d_sel, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
  std::vector<int32_t> prim((size_t)N);
  HIP_CHECK(hipMemcpy(prim.data(), c->d_sel_q + c->sel_cap, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
  for (int32_t i = 0; i < N; ++i) {
    const double* r = v.data() + (size_t)i * (2 * nx + 3);
    if (o->prim) o->prim[i] = prim[i];
    if (o->pair) o->pair[i] = q[i];
    if (o->distance) o->distance[i] = r[2 * nx];
    for (int k = 0; k < nx; ++k) {
      if (o->state) o->state[(int64_t)i * nx + k] = r[k];
      if (o->start) o->start[(int64_t)i * nx + k] = r[nx + k];
    }
    if (o->g) o->g[i] = r[2 * nx + 1];
    if (o->h) o->h[i] = r[2 * nx + 2];
  }
  return DP_OK;
}

int dp_expand_batch(dp_ctx* c, const double* nodes, const double* g_parent, int32_t N,
                    const dp_expand_params* p, void* stream) {
  int rc = dp_upload_nodes(c, nodes, g_parent, N);
  if (rc) return rc;
  rc = dp_expand(c, p, stream);
  if (rc) return rc;
  HIP_CHECK(hipStreamSynchronize(stream ? (hipStream_t)stream : c->stream));
  return DP_OK;
}

static int read_counters(dp_ctx* c, long long* h) {
  HIP_CHECK(hipSetDevice(c->device));
  HIP_CHECK(hipDeviceSynchronize());
  HIP_CHECK(hipMemcpy(h, c->d_counters, sizeof(long long) * DPC_COUNT, hipMemcpyDeviceToHost));
  return DP_OK;
}

int dp_get_counts(dp_ctx* c, dp_counts* o) {
  ARG_CHECK(c && o, "null argument");
  long long h[DPC_COUNT];
  int rc = read_counters(c, h);
  if (rc) return rc;
  // A quadrotor node whose cap did not fall inside its FP64 window: roll out
  // the deferred pairs in FP64 and resolve the batch again (no pair is
  // deferred the second time).
  const long long deferred = h[DPC_DEFERRED];
  if (h[DPC_WINDOW_OVF] > 0 && c->N > 0) {
    ++c->host_completions;
    if ((rc = dp_launch_quad_deferred(c, c->stream))) return rc;
    if ((rc = dp_launch_resolve_again(c, c->stream))) return rc;
    if ((rc = read_counters(c, h))) return rc;
    if (h[DPC_WINDOW_OVF] > 0) {
      dp_set_error("FP64 window overflow after the deferred pairs were rolled out");
      return DP_ERR_NOTCONV;
    }
  }
  // Resolution rounds run ahead blind; finish any node still undecided.
  if (h[DPC_NODES_LEFT] > 0 && c->N > 0) ++c->host_completions;
  for (int extra = 0; h[DPC_NODES_LEFT] > 0 && c->N > 0; ++extra) {
    if (extra > 1000) {
      dp_set_error("duplicate resolution did not converge");
      return DP_ERR_NOTCONV;
    }
    rc = dp_launch_more_rounds(c, 32, c->stream);
    if (rc) return rc;
    rc = read_counters(c, h);
    if (rc) return rc;
  }
  std::memset(o, 0, sizeof(*o));
  o->nodes = c->N;
  o->pairs = h[DPC_PAIRS];
  o->bounds = h[DPC_BOUNDS];
  o->collide = h[DPC_COLLIDE];
  o->free_ = h[DPC_FREE];
  o->surrogate_reject = h[DPC_SURR];
  o->novel = h[DPC_NOVEL];
  o->dup_visited = h[DPC_DUPV];
  o->dup_batch = h[DPC_DUPB];
  o->capped = h[DPC_CAPPED];
  o->goal_hits = h[DPC_GOAL];
  o->checked = h[DPC_CHECKED];
  o->recheck_fp64 = h[DPC_RECHECK];
  o->resolve_rounds = h[DPC_ROUNDS];
  o->conf_pool_used = h[DPC_CONF];
  o->fp64_deferred = deferred;
  o->constraint = h[DPC_CONSTRAINT];
  if (h[DPC_OVERFLOW]) {
    dp_set_error("pair capacity exceeded");
    return DP_ERR_CAPACITY;
  }
  if (h[DPC_ERRFLAG] & ~2ll) {
    dp_set_error("device error flag set");
    return DP_ERR_NOTCONV;
  }
  return DP_OK;
}

int dp_health(dp_ctx* c, dp_health_info* o) {
  ARG_CHECK(c && o, "null argument");
  HIP_CHECK(hipSetDevice(c->device));
  HIP_CHECK(hipDeviceSynchronize());
  long long st[DPS_COUNT];
  HIP_CHECK(hipMemcpy(st, c->d_sticky, sizeof(st), hipMemcpyDeviceToHost));
  std::memset(o, 0, sizeof(*o));
  o->batches = c->n_batches;
  o->wave_giveups = st[DPS_GIVEUP];
  o->window_overflows = st[DPS_WINDOW];
  o->host_completions = c->host_completions;
  return DP_OK;
}

int dp_debug_set_limits(dp_ctx* c, int32_t conf_slots, int64_t pool_entries, int32_t fp64_window) {
  ARG_CHECK(c, "null context");
  ARG_CHECK(conf_slots >= 0 && conf_slots <= DP_CONF_SLOTS, "conf_slots must be in [0, 64] (0: 64)");
  ARG_CHECK(pool_entries >= 0 && fp64_window >= 0, "negative limit");
  c->conf_slots = conf_slots > 0 ? conf_slots : DP_CONF_SLOTS;
  c->pool_limit = pool_entries;
  c->fp64_window = fp64_window;
  dp_bump_epoch();
  return DP_OK;
}

int dp_fetch_pairs(dp_ctx* c, dp_pairs_out* o) {
  ARG_CHECK(c && o, "null argument");
  dp_counts cnt;
  int rc = dp_get_counts(c, &cnt);
  if (rc) return rc;
  const int64_t n = cnt.pairs;
  if (n > o->capacity) {
    dp_set_error("output capacity smaller than the number of pairs");
    return DP_ERR_CAPACITY;
  }
  if (n == 0) return DP_OK;
  if (o->node) HIP_CHECK(hipMemcpy(o->node, c->d_pnode, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  if (o->prim) HIP_CHECK(hipMemcpy(o->prim, c->d_pprim, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  if (o->verdict) HIP_CHECK(hipMemcpy(o->verdict, c->d_verdict, n, hipMemcpyDeviceToHost));
  if (o->first_k) HIP_CHECK(hipMemcpy(o->first_k, c->d_firstk, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  if (o->goal_index) HIP_CHECK(hipMemcpy(o->goal_index, c->d_goal, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  std::vector<int8_t> vd((size_t)n), rs((size_t)n);
  HIP_CHECK(hipMemcpy(vd.data(), c->d_verdict, n, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(rs.data(), c->d_res, n, hipMemcpyDeviceToHost));
  if (o->resolution) std::memcpy(o->resolution, rs.data(), n);
  // A CAPPED pair comes after its node's loop ended (dbastar.cpp:916-918):
  // the reference computes no goal index, final state, g or h for it.
  for (int64_t q = 0; q < n; ++q)
    if (rs[q] == DP_R_CAPPED) vd[q] = -1;  // masked below like a non-FREE pair
  if (o->goal_index)
    for (int64_t q = 0; q < n; ++q)
      if (rs[q] == DP_R_CAPPED) o->goal_index[q] = -1;
  const double nan = NAN;
  if (o->final_state) {
    const int nx = c->desc.nx;
    std::vector<double> col((size_t)n);
    for (int k = 0; k < nx; ++k) {
      HIP_CHECK(hipMemcpy(col.data(), c->d_final + (size_t)k * c->pair_cap, sizeof(double) * n,
                          hipMemcpyDeviceToHost));
      for (int64_t q = 0; q < n; ++q) o->final_state[q * nx + k] = vd[q] == DP_V_FREE ? col[q] : nan;
    }
  }
  if (o->g) {
    HIP_CHECK(hipMemcpy(o->g, c->d_g, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < n; ++q)
      if (vd[q] != DP_V_FREE) o->g[q] = nan;
  }
  if (o->h) {
    HIP_CHECK(hipMemcpy(o->h, c->d_h, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < n; ++q)
      if (vd[q] != DP_V_FREE) o->h[q] = nan;
  }
  return DP_OK;
}

int dp_visited_append_novel(dp_ctx* c) {
  ARG_CHECK(c, "null context");
  dp_counts cnt;
  int rc = dp_get_counts(c, &cnt);
  if (rc) return rc;
  if (cnt.novel == 0) return DP_OK;
  rc = ensure_visited_cap(c, c->nv + cnt.novel);
  if (rc) return rc;
  rc = dp_launch_visited_append_novel(c, cnt.novel, c->stream);
  if (rc) return rc;
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return DP_OK;
}


int dp_last_timing(dp_ctx* c, float* ms5) {
  ARG_CHECK(c && ms5, "null argument");
  if (!c->timing_valid) {
    dp_set_error("no expansion has been timed");
    return DP_ERR_STATE;
  }
  HIP_CHECK(hipEventSynchronize(c->ev[3]));
  HIP_CHECK(hipEventElapsedTime(&ms5[0], c->ev[0], c->ev[4]));
  HIP_CHECK(hipEventElapsedTime(&ms5[1], c->ev[1], c->ev[2]));
  HIP_CHECK(hipEventElapsedTime(&ms5[2], c->ev[2], c->ev[3]));
  HIP_CHECK(hipEventElapsedTime(&ms5[3], c->ev[0], c->ev[3]));
  HIP_CHECK(hipEventElapsedTime(&ms5[4], c->ev[4], c->ev[1]));
  return DP_OK;
}

int dp_last_timing_detail(dp_ctx* c, float* ms7) {
  ARG_CHECK(c && ms7, "null argument");
  if (!c->timing_valid) {
    dp_set_error("no expansion has been timed");
    return DP_ERR_STATE;
  }
  HIP_CHECK(hipEventSynchronize(c->ev[3]));
  const int from[7] = {0, 4, 1, 5, 6, 2, 7}, to[7] = {4, 1, 5, 6, 2, 7, 3};
  for (int k = 0; k < 7; ++k) HIP_CHECK(hipEventElapsedTime(&ms7[k], c->ev[from[k]], c->ev[to[k]]));
  return DP_OK;
}

int dp_set_tiles(dp_ctx* c, const dp_tile_config* t) {
  ARG_CHECK(c && t, "null argument");
  const int l = t->lanes_per_pair, q = t->quad64_threads;
  ARG_CHECK(l == 0 || l == 2 || l == 4 || l == 8 || l == 16 || l == 32 || l == 64,
            "lanes_per_pair must be 0, 2, 4, 8, 16, 32 or 64");
  ARG_CHECK(q == 64 || q == 128 || q == 256, "quad64_threads must be 64, 128 or 256");
  ARG_CHECK(t->check_blocks >= 1 && t->check_blocks <= 65535 && t->quad32_blocks >= 1 &&
                t->quad32_blocks <= 65535 && t->quad64_blocks >= 1 && t->quad64_blocks <= 65535,
            "block counts must be in [1, 65535]");
  c->tiles = *t;
  dp_bump_epoch();
  return DP_OK;
}

int dp_get_tiles(const dp_ctx* c, dp_tile_config* t) {
  ARG_CHECK(c && t, "null argument");
  *t = c->tiles;
  return DP_OK;
}

int dp_autotune(dp_ctx* c, const dp_expand_params* p, dp_tile_config* best, dp_tile_config* trials,
                int32_t* n_trials) {
  ARG_CHECK(c && p && best, "null argument");
  std::vector<dp_tile_config> cand;
  const dp_tile_config base = c->tiles;
  if (c->desc.system == DP_SYS_QUADROTOR) {
    for (int q32 : {512, 1024, 2048, 4096})
      for (int t64 : {64, 128, 256})
        for (int b64 : {4096, 8192, 16384}) {
          dp_tile_config t = base;
          t.quad32_blocks = q32;
          t.quad64_threads = t64;
          t.quad64_blocks = b64;
          cand.push_back(t);
        }
  } else {
    for (int l : {2, 4, 8, 16, 32, 64})
      for (int b : {1024, 2048, 4096, 8192}) {
        dp_tile_config t = base;
        t.lanes_per_pair = l;
        t.check_blocks = b;
        cand.push_back(t);
      }
  }
  int best_i = -1;
  for (size_t i = 0; i < cand.size(); ++i) {
    int rc = dp_set_tiles(c, &cand[i]);
    if (rc) return rc;
    float ms[5], runs[3];
    for (int r = -1; r < 3; ++r) {
      if ((rc = dp_expand(c, p, nullptr))) return rc;
      if ((rc = dp_last_timing(c, ms))) return rc;
      if (r >= 0) runs[r] = ms[1];
    }
    std::sort(runs, runs + 3);
    cand[i].check_ms = runs[1];
    if (best_i < 0 || cand[i].check_ms < cand[best_i].check_ms) best_i = (int)i;
  }
  c->tiles = cand[best_i];
  dp_bump_epoch();
  *best = cand[best_i];
  if (trials && n_trials) {
    const int32_t n = (int32_t)std::min<size_t>(cand.size(), 64);
    for (int32_t i = 0; i < n; ++i) trials[i] = cand[i];
    *n_trials = n;
  }
  // the counters of the last timed run stay valid; finish its resolution
  dp_counts cnt;
  return dp_get_counts(c, &cnt);
}

int dp_debug_resolve_trace(dp_ctx* c, int32_t enable, uint64_t* out, int64_t n) {
  ARG_CHECK(c && n >= 0, "invalid argument");
  HIP_CHECK(hipSetDevice(c->device));
  HIP_CHECK(hipDeviceSynchronize());
  if (out && n > 0) {
    ARG_CHECK(c->d_trace && n <= c->trace_cap / DP_TRACE_FIELDS, "no trace of that size");
    HIP_CHECK(hipMemcpy(out, c->d_trace, sizeof(uint64_t) * DP_TRACE_FIELDS * n, hipMemcpyDeviceToHost));
    int khz = 0;
    HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
    out[DP_TRACE_FIELDS * n] = (uint64_t)khz;
  }
  if (enable > 0) {
    const int64_t need = DP_TRACE_FIELDS * (int64_t)std::max(c->N_cap, 1);
    if (need > c->trace_cap) {
      int rc = dev_alloc(&c->d_trace, need);
      if (rc) return rc;
      c->trace_cap = need;
    }
  } else if (enable < 0) {
    dev_free(&c->d_trace);
    c->trace_cap = 0;
  }
  return DP_OK;
}

}  // extern "C"
