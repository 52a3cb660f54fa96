// k_expand.hip -- gfx950 kernels of the batched db-A* expansion.
//
// Pipeline of one dp_expand (all on one stream, no host round trip):
//   k_app_count       one wave per node: applicability query nearestR(c, alpha*delta)
//                     over primitive start states (Expander::expand_lazy,
//                     dbastar.hpp:270-312), brute force with exact early
//                     rejections, one bit mask per 64 primitives, counted
//   k_scan            exclusive scan -> canonical pair offsets per node
//   k_app_fill        the masks as the (node, primitive) pair list
//   k_check_planar    LPP lanes per pair, one lane per primitive state:
//                     translate (transform_primitive), is_state_valid on every
//                     state, swept-box cull of LDS-staged obstacles, SAT
//                     per state, first offending index by ballot
//                     (check_lazy_trajectory, dbastar.cpp:391-490), check_goal
//                     (:492-512), g/h (:833-847), duplicate test against the
//                     visited grid (:850-854)
//   k_grid.hip        contender grid, earlier-conflict lists and the round-
//                     based duplicate / branching-cap resolution (:856-918)
#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>

#include <cmath>

#include "dp_internal.h"

namespace {

constexpr int kWaves = 4;  // waves per workgroup of the check kernels

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------------
// applicability (T_m->nearestR over primitive start states)
// ------------------------------------------------------------------------
// Exact rejections ahead of dp_within for the applicability scan: each test
// shows d(a, b) > r from one metric term (the terms are non-negative and
// IEEE addition is monotone), reading a few components only.
struct app_bounds {
  double T_p, T_v, T_w;  // squared-sum bounds of dp_pos_bound's form for the position,
                         // velocity and rate terms (quadrotor; INFINITY when unused)
  double c_q;            // quadrotor: |q_a . q_b| < c_q implies w1 * angle > r
};

template <int NX>
__device__ __forceinline__ bool app_reject(const dp_dev_sys& S, const app_bounds& B, const double* a,
                                           const double* start, int32_t P, int32_t j, double r) {
  if constexpr (NX == 13) {
    double dot = a[3] * start[(int64_t)3 * P + j];
    dot = dot + a[4] * start[(int64_t)4 * P + j];
    dot = dot + a[5] * start[(int64_t)5 * P + j];
    dot = dot + a[6] * start[(int64_t)6 * P + j];
    if (fabs(dot) < B.c_q) return true;
    double s = 0.0;
#pragma unroll
    for (int c = 10; c < 13; ++c) {
      const double d = a[c] - start[(int64_t)c * P + j];
      s = c == 10 ? d * d : s + d * d;
    }
    return s > B.T_w;
  } else {
    // the heading term (unicycles, car): w1 * so2 alone
    return S.w[1] * dp_so2(a[2], start[(int64_t)2 * P + j]) > r;
  }
}

// kAppParts waves per node, each over a share of the primitives, 64 per
// step: the applicable ones are recorded as a bit mask per 64 primitives
// (mask[i * words + j / 64]) and counted into app[i] (cleared before); the
// fill pass turns the masks into the (node, primitive) pair list without
// evaluating the metric again.
constexpr int kAppParts = 4;

// With a pair order (order != nullptr, [rows][P]) bit r of node i's mask is
// primitive order[(i % rows) * P + r]: the pairs come out in that order.
template <int NX>
__global__ __launch_bounds__(64) void k_app_count(const dp_dev_sys* __restrict__ Sp, app_bounds B,
                                                  const double* __restrict__ nodes, int32_t Ncap, int32_t N,
                                                  const double* __restrict__ start, int32_t P, double r_app,
                                                  const int32_t* __restrict__ order, int32_t rows,
                                                  int32_t* __restrict__ app, uint64_t* __restrict__ mask) {
  const dp_dev_sys& S = *Sp;
  const int32_t i = blockIdx.x / kAppParts, part = blockIdx.x % kAppParts;
  if (i >= N) return;
  const int lane = threadIdx.x;
  const int32_t words = (P + 63) / 64;
  const int32_t w0 = (int32_t)((int64_t)words * part / kAppParts), w1 = (int32_t)((int64_t)words * (part + 1) / kAppParts);
  const int32_t* row = order ? order + (int64_t)(i % rows) * P : nullptr;
  double a[NX];
#pragma unroll
  for (int c = 0; c < NX; ++c) a[c] = dp_is_offset_dim(S, c) ? 0.0 : nodes[(int64_t)c * Ncap + i];
  int32_t cnt = 0;
  for (int32_t j0 = 64 * w0; j0 < P && j0 < 64 * w1; j0 += 64) {
    const int32_t r = j0 + lane;
    const int32_t j = row && r < P ? row[r] : r;
    bool hit = false;
    if (r < P && !app_reject<NX>(S, B, a, start, P, j, r_app)) {
      double b[NX];
#pragma unroll
      for (int c = 0; c < NX; ++c) b[c] = start[(int64_t)c * P + j];
      hit = dp_within(S, a, b, r_app);
    }
    const uint64_t m = __ballot(hit);
    if (lane == 0) mask[(int64_t)i * words + j0 / 64] = m;
    cnt += __popcll(m);
  }
  if (lane == 0 && cnt) atomicAdd(&app[i], cnt);
}

// The same query for kAppNodes nodes per wave (a shared pair order or
// none): each lane loads its primitive's key once per 64-primitive chunk and
// tests it against every node of the group, so the keys stream from L2 once
// per group instead of once per node.  The exact early rejections run for
// all (node, primitive) combinations; the combinations they leave go to a
// wave-local queue and the full metric runs on them 64 at a time, packed,
// instead of on every chunk that has at least one such lane.  Bit r of node
// i's mask word w is set exactly when k_app_count would set it.
constexpr int kAppNodes = 16;

// Exact rejections of the blocked query, evaluated in FP32 on FP32 copies of
// the keys and node states, with margins that cover FP32 rounding (about
// 1e-6 relative on the quantities compared, against 1e-4 of margin): a pair
// rejected here is rejected by the FP64 metric too.  Quadrotor: the attitude
// term is w1 * acos(|q_a . q_b|), and for unit quaternions acos(|q_a . q_b|)
// >= |q_a - sign(q_a . q_b) q_b| (the chord 2 sin(theta / 2) <= theta); every
// other term is non-negative, so w1 * chord + w3 * |w_a - w_b| above the
// radius proves d > r.  That leaves about 0.3% more pairs than are
// applicable for the FP64 metric, against about 3x with the attitude bound
// alone.  Planar systems: the heading term alone.
struct app_bounds32 {
  float c_q;      // quadrotor: |dot| below it rejects
  float T_w;      // quadrotor: squared rate difference above it rejects
  float w1, w3;   // metric weights of the attitude / heading and rate terms
  float R;        // quadrotor: chord bound above it rejects; planar: heading term above it rejects
};

// Stage 1, for every (node, primitive) of a chunk: quadrotor |dot| >= c_q;
// planar systems: the heading term.
template <int NX>
__device__ __forceinline__ bool app_pass1(const app_bounds32& B, const float* a, const float* k) {
  if constexpr (NX == 13) {
    float dot = a[3] * k[3];
    dot = dot + a[4] * k[4];
    dot = dot + a[5] * k[5];
    dot = dot + a[6] * k[6];
    return !(fabsf(dot) < B.c_q);
  } else {
    float d = fabsf(a[2] - k[2]);
    d = d > 3.14159265f ? 6.28318531f - d : d;
    return !(B.w1 * d > B.R);
  }
}

// Stage 2, quadrotor only, on the stage-1 survivors (packed): the chord and
// rate bound.
__device__ __forceinline__ bool app_pass2_quad(const app_bounds32& B, const float* a, const float* k) {
  float dot = a[3] * k[3];
  dot = dot + a[4] * k[4];
  dot = dot + a[5] * k[5];
  dot = dot + a[6] * k[6];
  const float sg = dot < 0.0f ? -1.0f : 1.0f;
  // sums of squared differences with their FP32 error bounds: each term
  // (a - k)^2 is off by at most 8e-7 (a^2 + k^2) (inputs rounded to FP32,
  // one subtraction, one product, the sum); 2e-6 (a^2 + k^2) is taken off
  float c2 = 0.0f, s = 0.0f, ec = 0.0f, es = 0.0f;
#pragma unroll
  for (int c = 3; c < 7; ++c) {
    const float d = a[c] - sg * k[c];
    c2 = c2 + d * d;
    ec = ec + (a[c] * a[c] + k[c] * k[c]);
  }
#pragma unroll
  for (int c = 10; c < 13; ++c) {
    const float d = a[c] - k[c];
    s = s + d * d;
    es = es + (a[c] * a[c] + k[c] * k[c]);
  }
  const float c2_lo = fmaxf(c2 - 2e-6f * ec, 0.0f), s_lo = fmaxf(s - 2e-6f * es, 0.0f);
  if (s_lo > B.T_w) return false;
  return !(B.w1 * sqrtf(c2_lo) + B.w3 * sqrtf(s_lo) > B.R);
}

template <int NX>
__global__ __launch_bounds__(64) void k_app_blk(const dp_dev_sys* __restrict__ Sp, app_bounds32 B,
                                                const double* __restrict__ nodes, int32_t Ncap, int32_t N,
                                                const double* __restrict__ keys, const float* __restrict__ keys32,
                                                int32_t P, double r_app, const int32_t* __restrict__ row,
                                                int32_t parts, int32_t* __restrict__ app,
                                                uint64_t* __restrict__ mask) {
  __shared__ double s_a[kAppNodes * NX];
  __shared__ float s_a32[kAppNodes * NX];
  __shared__ float s_k[NX * 64];
  __shared__ uint16_t s_q[kAppNodes * 64];
  __shared__ uint16_t s_q2[kAppNodes * 64];
  __shared__ unsigned long long s_m[kAppNodes];
  const dp_dev_sys& S = *Sp;
  const int32_t grp = blockIdx.x / parts, part = blockIdx.x % parts;
  const int32_t i0 = grp * kAppNodes;
  if (i0 >= N) return;
  const int32_t nb = N - i0 < kAppNodes ? N - i0 : kAppNodes;
  const int lane = threadIdx.x;
  const uint64_t lt = (1ull << lane) - 1ull;
  const int32_t words = (P + 63) / 64;
  const int32_t w0 = (int32_t)((int64_t)words * part / parts), w1 = (int32_t)((int64_t)words * (part + 1) / parts);
  for (int t = lane; t < kAppNodes * NX; t += 64) {
    const int b = t / NX, c = t % NX;
    s_a[t] = b < nb && !dp_is_offset_dim(S, c) ? nodes[(int64_t)c * Ncap + i0 + b] : 0.0;
    s_a32[t] = (float)s_a[t];
  }
  if (lane < kAppNodes) s_m[lane] = 0ull;
  wave_lds_sync();
  int32_t cnt = 0;  // lane b < nb: node i0 + b
  for (int32_t w = w0; w < w1; ++w) {
    const int32_t r = w * 64 + lane;
    // every lane loads (past P, the last primitive's key, then unused): a
    // load under a condition would be waited for on its own, component
    // after component
    const int32_t rc = r < P ? r : P - 1;
    const int32_t j = row ? row[rc] : rc;
    float k[NX];
#pragma unroll
    for (int c = 0; c < NX; ++c) k[c] = keys32[(int64_t)c * P + j];
#pragma unroll
    for (int c = 0; c < NX; ++c) {
      k[c] = r < P ? k[c] : 0.0f;
      s_k[c * 64 + lane] = k[c];
    }
    // stage 1 for every node of the group; survivors queued
    int qn = 0;
    for (int b = 0; b < nb; ++b) {
      const bool pass = r < P && app_pass1<NX>(B, s_a32 + b * NX, k);
      const uint64_t m = __ballot(pass);
      if (pass) s_q[qn + __popcll(m & lt)] = (uint16_t)((b << 6) | lane);
      qn += __popcll(m);
    }
    wave_lds_sync();
    const uint16_t* q = s_q;
    if constexpr (NX == 13) {
      // stage 2 on the packed survivors
      int qn2 = 0;
      for (int e0 = 0; e0 < qn; e0 += 64) {
        const int e = e0 + lane;
        bool pass = false;
        int ent = 0;
        if (e < qn) {
          ent = s_q[e];
          const int b = ent >> 6, lj = ent & 63;
          float kk[NX];
#pragma unroll
          for (int c = 0; c < NX; ++c) kk[c] = s_k[c * 64 + lj];
          pass = app_pass2_quad(B, s_a32 + b * NX, kk);
        }
        const uint64_t m = __ballot(pass);
        if (pass) s_q2[qn2 + __popcll(m & lt)] = (uint16_t)ent;
        qn2 += __popcll(m);
      }
      wave_lds_sync();
      q = s_q2;
      qn = qn2;
    }
    // the full metric on the survivors, 64 at a time
    for (int e0 = 0; e0 < qn; e0 += 64) {
      const int e = e0 + lane;
      if (e < qn) {
        const int ent = q[e];
        const int b = ent >> 6, lj = ent & 63;
        const int32_t jj = row ? row[w * 64 + lj] : w * 64 + lj;
        double kk[NX];
#pragma unroll
        for (int c = 0; c < NX; ++c) kk[c] = keys[(int64_t)c * P + jj];
        if (dp_within(S, s_a + b * NX, kk, r_app)) atomicOr(&s_m[b], 1ull << lj);
      }
    }
    wave_lds_sync();
    if (lane < nb) {
      const unsigned long long m = s_m[lane];
      mask[(int64_t)(i0 + lane) * words + w] = m;
      cnt += __popcll(m);
      s_m[lane] = 0ull;
    }
    wave_lds_sync();
  }
  if (lane < nb && cnt) atomicAdd(&app[i0 + lane], cnt);
}


// The masks as the (node, primitive) pair list, one wave per node, 64 mask
// words at a time: each lane expands its own word's set bits into LDS at the
// word's offset (an exclusive scan of the counts across the lanes), then the
// wave copies the batch out with full-width stores.
template <bool kRow>
__global__ __launch_bounds__(64) void k_app_fill(int32_t N, int32_t P, const int32_t* __restrict__ app,
                                                 const uint64_t* __restrict__ mask, const int32_t* __restrict__ order,
                                                 int32_t rows, int32_t* __restrict__ pnode,
                                                 int32_t* __restrict__ pprim, int64_t pair_cap) {
  __shared__ int32_t s_p[64 * 64];
  const int32_t i = blockIdx.x;
  if (i >= N) return;
  const int lane = threadIdx.x;
  const int32_t words = (P + 63) / 64;
  const int32_t* row = kRow ? order + (int64_t)(i % rows) * P : nullptr;
  int64_t q0 = app[i];
  for (int32_t wb = 0; wb < words; wb += 64) {
    const int32_t wl = wb + lane;
    uint64_t m = wl < words ? mask[(int64_t)i * words + wl] : 0ull;
    int32_t incl = __popcll(m);
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t v = __shfl_up(incl, o);
      if (lane >= o) incl += v;
    }
    int32_t pos = incl - __popcll(m);
    while (m) {
      s_p[pos++] = wl * 64 + (__ffsll((long long)m) - 1);
      m &= m - 1ull;
    }
    wave_lds_sync();
    const int32_t total = __builtin_amdgcn_readlane(incl, 63);
    for (int32_t k = lane; k < total; k += 64) {
      const int64_t q = q0 + k;
      if (q < pair_cap) {
        const int32_t r = s_p[k];
        pnode[q] = i;
        if constexpr (kRow) pprim[q] = row[r];
        else pprim[q] = r;
      }
    }
    q0 += total;
    wave_lds_sync();
  }
}

}  // namespace

// Arguments of the check kernels (passed by value in the kernarg segment).
struct dp_check_args {
  const dp_dev_sys* sys;
  const double* nodes;
  int32_t Ncap;
  const double* gpar;
  const int32_t* off;
  const double* st;
  int32_t S;
  const double* geo;
  int32_t nbox;
  const double* sweep;
  const double* srange;  // [2][NX][P] per-primitive min / max of each component
  int32_t P;
  const int32_t* goalidx;
  int32_t num_goal, check_goal;
  const double* obs;
  int32_t nobs;
  const int16_t* node_obs;   // [Ncap][DP_NODE_OBS] ascending obstacle indices near each node
  const int32_t* node_nobs;  // [Ncap] their count, -1 when the list overflowed (scan all)
  const double* vis;
  int64_t vstride;
  int64_t nv;
  const int32_t* vstart;
  dp_grid_geom grid;
  const int32_t* pnode;
  const int32_t* pprim;
  const int32_t* total;
  int8_t* verdict;
  int32_t* firstk;
  int32_t* goal_out;
  int8_t* res;
  int8_t* flags;
  double* fin;
  int64_t pcap;
  double* gout;
  double* hout;
  double goal[DP_MAX_NX];
  double goal_bound, r_dup, cdf;
  int32_t surrogate;  // pairs already marked DP_V_SURROGATE_REJECT are skipped
  long long* counters;
  // tdbastar's reverse search: placement offset = node - last position of
  // the primitive, new state = first placed state
  int32_t backward;
  const double* last;  // [nx][P]
  // tdbastar constraints: times [ncon], then states [ncon][nx]
  int32_t ncon;
  const double* con;
  double con_delta;
  // dbrrt selection: no goal check or duplicate test, resolution DP_R_NONE
  int32_t select;
};

namespace {

template <int NX>
__device__ __forceinline__ void load_translated(const dp_check_args& a, int64_t sidx, double ox, double oy, double* s) {
  s[0] = a.st[sidx] + ox;
  s[1] = a.st[(int64_t)a.S + sidx] + oy;
#pragma unroll
  for (int c = 2; c < NX; ++c) s[c] = a.st[(int64_t)c * a.S + sidx];
}


// Per frontier node, the obstacles that can touch any of its primitives: the
// swept box of every primitive lies inside the union box `reach`, and the
// pair test below uses the same rounding form with the larger box, so every
// obstacle a pair's swept-box cull keeps is on its node's list.  One wave per
// node; lists hold ascending indices.
__global__ __launch_bounds__(256) void k_node_obs(const double* nodes, int32_t Ncap, int32_t N, const double* obs,
                                                  int32_t nobs, double r0, double r1, double r2, double r3,
                                                  int16_t* node_obs, int32_t* node_nobs) {
  const int lane = threadIdx.x % 64;
  for (int64_t i = (int64_t)blockIdx.x * kWaves + threadIdx.x / 64; i < N; i += (int64_t)gridDim.x * kWaves) {
    const double ox = nodes[i], oy = nodes[(int64_t)Ncap + i];
    const double bx0 = (r0 + ox) - DP_SWEEP_PAD, bx1 = (r1 + ox) + DP_SWEEP_PAD;
    const double by0 = (r2 + oy) - DP_SWEEP_PAD, by1 = (r3 + oy) + DP_SWEEP_PAD;
    int cnt = 0;
    for (int ob = 0; ob < nobs; ob += 64) {
      const int o = ob + lane;
      bool hit = false;
      if (o < nobs) {
        const double* b = obs + 4 * o;
        hit = !(b[0] + b[2] < bx0 || b[0] - b[2] > bx1 || b[1] + b[3] < by0 || b[1] - b[3] > by1);
      }
      const uint64_t m = __ballot(hit);
      const int at = cnt + __popcll(m & ((1ull << lane) - 1ull));
      if (hit && at < DP_NODE_OBS) node_obs[i * DP_NODE_OBS + at] = (int16_t)o;
      cnt += __popcll(m);
    }
    if (lane == 0) node_nobs[i] = cnt <= DP_NODE_OBS ? cnt : -1;
  }
}

// Planar translate systems (unicycle1/2, car+trailer), FP64.
template <int NX, int LPP>
__global__ __launch_bounds__(256) void k_check_planar(dp_check_args a) {
  static_assert(LPP >= 2 && 64 % LPP == 0, "a pair takes 2 to 64 lanes of a wave");
  extern __shared__ double smem[];
  // the system constants in LDS: read from global memory after the loop's
  // stores, each would be a load the wave waits for
  __shared__ dp_dev_sys s_sys;
  const dp_dev_sys& S = s_sys;
  const int nobs = a.nobs;
  if ((int64_t)blockIdx.x * kWaves * (64 / LPP) >= (int64_t)*a.total) return;  // uniform per block
  double* s_obs = smem;  // [nobs][4]
  int16_t* s_cand = (int16_t*)(smem + 4 * nobs);
  for (int t = threadIdx.x; t < 4 * nobs; t += blockDim.x) s_obs[t] = a.obs[t];
  static_assert(sizeof(dp_dev_sys) % sizeof(double) == 0, "copied as doubles");
  for (int t = threadIdx.x; t < (int)(sizeof(dp_dev_sys) / sizeof(double)); t += blockDim.x)
    reinterpret_cast<double*>(&s_sys)[t] = reinterpret_cast<const double*>(a.sys)[t];
  // the workgroup's counts, added to the batch counters once per workgroup
  __shared__ unsigned long long s_cnt[5];
  if (threadIdx.x < 5) s_cnt[threadIdx.x] = 0ull;
  __syncthreads();

  constexpr int SEG = 64 / LPP;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
  const int seg = lane / LPP, k0 = lane % LPP;
  const int shift = seg * LPP;
  const uint64_t segbits = LPP == 64 ? ~0ull : ((1ull << LPP) - 1ull);
  const uint64_t ltmask = (1ull << k0) - 1ull;
  int16_t* cand = s_cand + (int64_t)(wave * SEG + seg) * nobs;
  const int32_t total = *a.total;
  const double Tdup = dp_pos_bound(S, a.r_dup);
  long long c_b = 0, c_c = 0, c_f = 0, c_k = 0, c_x = 0;
  // bounds and footprint held in registers: read inside the loops below
  // (after the loop's stores) every bound would be a load and a wait
  double blo[NX], bhi[NX];
#pragma unroll
  for (int c = 0; c < NX; ++c) {
    blo[c] = S.lo[c];
    bhi[c] = S.hi[c];
  }
  const double body_hl = S.body_hl, body_hw = S.body_hw, trailer_hl = S.trailer_hl, trailer_hw = S.trailer_hw;

  for (int64_t q0 = ((int64_t)blockIdx.x * kWaves + wave) * SEG; q0 < total;
       q0 += (int64_t)gridDim.x * kWaves * SEG) {
    const int64_t q = q0 + seg;
    const bool vp = q < total && !(a.surrogate && a.verdict[q] == DP_V_SURROGATE_REJECT);
    int32_t i = 0, j = 0, st0 = 0, L1 = 0;
    double ox = 0.0, oy = 0.0;
    if (vp) {
      i = a.pnode[q];
      j = a.pprim[q];
      st0 = a.off[j];
      L1 = a.off[j + 1] - st0;
      ox = a.nodes[i];
      oy = a.nodes[(int64_t)a.Ncap + i];
      if (a.backward) {  // offset - motion->traj.states.back().head(2), tdbastar.hpp:193-196
        ox = ox - a.last[j];
        oy = oy - a.last[(int64_t)a.P + j];
      }
    }
    // ---- is_state_valid on every translated state -----------------------
    // First from the primitive's component ranges: with the translation
    // rounded the same way, (min + offset) >= lo and (max + offset) <= hi
    // bound every translated state (rounding is monotone); other components
    // compare their range directly.  Only pairs not cleared this way scan
    // their states for the first offending index.
    bool clear = true;
    if (vp) {
#pragma unroll
      for (int c = 0; c < NX; ++c) {
        const double o = c == 0 ? ox : (c == 1 ? oy : 0.0);
        const double mn = a.srange[(int64_t)c * a.P + j], mx = a.srange[((int64_t)NX + c) * a.P + j];
        const double tmn = c < 2 ? mn + o : mn, tmx = c < 2 ? mx + o : mx;
        clear = clear & (tmn >= blo[c]) & (tmx <= bhi[c]);
      }
    }
    int first_bad = INT_MAX;
    for (int kb = 0; __ballot(vp && !clear && kb < L1) != 0; kb += LPP) {
      const int k = kb + k0;
      bool bad = false;
      if (vp && !clear && k < L1) {
        const int64_t sidx = (int64_t)st0 + k;
        double s[NX];
        load_translated<NX>(a, sidx, ox, oy, s);
#pragma unroll
        for (int c = 0; c < NX; ++c) bad = bad | !((s[c] >= blo[c]) & (s[c] <= bhi[c]));
      }
      const uint64_t m = (__ballot(bad) >> shift) & segbits;
      if (first_bad == INT_MAX && m) first_bad = kb + __ffsll((long long)m) - 1;
    }
    int8_t vd = first_bad != INT_MAX ? DP_V_BOUNDS : DP_V_FREE;
    const bool need_col = vp && vd == DP_V_FREE;

    // ---- collision: swept-box cull, then SAT per state ------------------
    int first_col = INT_MAX;
    if (__ballot(need_col)) {
      double bx0 = 0, bx1 = 0, by0 = 0, by1 = 0;
      if (need_col) {
        bx0 = (a.sweep[j] + ox) - DP_SWEEP_PAD;
        bx1 = (a.sweep[(int64_t)a.P + j] + ox) + DP_SWEEP_PAD;
        by0 = (a.sweep[2 * (int64_t)a.P + j] + oy) - DP_SWEEP_PAD;
        by1 = (a.sweep[3 * (int64_t)a.P + j] + oy) + DP_SWEEP_PAD;
      }
      // scan the node's obstacle list (all obstacles when it overflowed)
      int nlist = 0;
      const int16_t* list = nullptr;
      if (need_col) {
        nlist = a.node_nobs[i];
        if (nlist >= 0) list = a.node_obs + (int64_t)i * DP_NODE_OBS;
        else nlist = nobs;
      }
      int ncand = 0;
      for (int ob = 0; __ballot(need_col && ob < nlist) != 0; ob += LPP) {
        const int t = ob + k0;
        bool hit = false;
        int o = 0;
        if (need_col && t < nlist) {
          o = list ? list[t] : t;
          const double* b = s_obs + 4 * o;
          hit = !(b[0] + b[2] < bx0 || b[0] - b[2] > bx1 || b[1] + b[3] < by0 || b[1] - b[3] > by1);
        }
        const uint64_t m = (__ballot(hit) >> shift) & segbits;
        if (hit) cand[ncand + __popcll(m & ltmask)] = (int16_t)o;
        ncand += __popcll(m);
      }
      wave_lds_sync();
      for (int kb = 0; __ballot(need_col && ncand > 0 && kb < L1) != 0; kb += LPP) {
        const int k = kb + k0;
        bool col = false;
        if (need_col && ncand > 0 && k < L1 && first_col == INT_MAX) {
          const int64_t sidx = (int64_t)st0 + k;
          const double x = a.st[sidx] + ox;
          const double y = a.st[(int64_t)a.S + sidx] + oy;
          for (int bi = 0; bi < a.nbox && !col; ++bi) {
            const double c = a.geo[(int64_t)(4 * bi + 2) * a.S + sidx];
            const double s = a.geo[(int64_t)(4 * bi + 3) * a.S + sidx];
            double cx = x, cy = y, hl = body_hl, hw = body_hw;
            if (bi > 0) {
              cx = x + a.geo[(int64_t)(4 * bi) * a.S + sidx];
              cy = y + a.geo[(int64_t)(4 * bi + 1) * a.S + sidx];
              hl = trailer_hl;
              hw = trailer_hw;
            }
            for (int t = 0; t < ncand; ++t) {
              const double* b = s_obs + 4 * cand[t];
              if (dp_obb_aabb(cx, cy, c, s, hl, hw, b[0], b[1], b[2], b[3])) {
                col = true;
                break;
              }
            }
          }
        }
        const uint64_t m = (__ballot(col) >> shift) & segbits;
        if (first_col == INT_MAX && m) first_col = kb + __ffsll((long long)m) - 1;
      }
      wave_lds_sync();  // candidate list is rewritten by the next pair group
      if (need_col && first_col != INT_MAX) vd = DP_V_COLLIDE;
    }

    // ---- tdbastar constraints on a valid pair (tdbastar.cpp:408-443) -----
    // LPP constraints at a time; the first violated one in list order wins.
    if (a.ncon > 0 && __ballot(vp && vd == DP_V_FREE)) {
      const bool nc = vp && vd == DP_V_FREE;
      bool reaches = false;
      float gf = 0.0f;
      if (nc) {
        double s[NX];
        load_translated<NX>(a, (int64_t)st0 + (a.backward ? L1 - 1 : 0), ox, oy, s);
        reaches = dp_distance(S, s, a.goal) <= a.con_delta;
        gf = (float)a.gpar[i];  // best_node_gScore is passed as float
      }
      int viol_k = -1;
      for (int cb = 0; __ballot(nc && viol_k < 0 && cb < a.ncon) != 0; cb += LPP) {
        const int ci = cb + k0;
        bool viol = false;
        int k = -1;
        if (nc && viol_k < 0 && ci < a.ncon) {
          const float toff = (float)(a.con[ci] - (double)gf);
          const int ti = (int)lround((double)toff / S.dt);
          if (reaches && ti >= L1 - 1) k = L1 - 1;
          if (ti >= 0 && ti < L1 - 1) k = ti;
          if (k >= 0) {
            double s[NX], cs[NX];
            load_translated<NX>(a, (int64_t)st0 + k, ox, oy, s);
#pragma unroll
            for (int c = 0; c < NX; ++c) cs[c] = a.con[a.ncon + (int64_t)ci * NX + c];
            viol = dp_distance(S, s, cs) <= a.con_delta;
          }
        }
        const uint64_t m = (__ballot(viol) >> shift) & segbits;
        const int src = m ? shift + __ffsll((long long)m) - 1 : lane;
        const int kk = __shfl(k, src);
        if (nc && viol_k < 0 && m) viol_k = kk;
      }
      if (nc && viol_k >= 0) {
        vd = DP_V_CONSTRAINT;
        first_col = viol_k;
      }
    }

    // ---- FREE: check_goal, g, h, duplicate vs visited --------------------
    const bool fr = vp && vd == DP_V_FREE;
    int chosen = -1;
    if (a.check_goal && !a.select && __ballot(fr)) {
      // goal indices LPP at a time; the first in index order that hits wins
      for (int gb = 0; gb < a.num_goal; gb += LPP) {
        const int gk = gb + k0;
        bool hit = false;
        if (fr && chosen == -1 && gk < a.num_goal) {
          const int idx = a.goalidx[(int64_t)gk * a.P + j];
          double s[NX];
          load_translated<NX>(a, (int64_t)st0 + idx, ox, oy, s);
          hit = dp_within(S, s, a.goal, a.goal_bound);
        }
        const uint64_t m = (__ballot(hit) >> shift) & segbits;
        if (fr && chosen == -1 && m) chosen = a.goalidx[(int64_t)(gb + __ffsll((long long)m) - 1) * a.P + j];
      }
    }
    double xf[NX];
    double gq = 0.0, hq = 0.0;
    if (fr) {
      // the new node's state: the last placed state, or the first for the
      // reverse search (check_goal, tdbastar.cpp:453-458)
      const int fi = chosen != -1 ? chosen : (a.backward ? 0 : L1 - 1);
      load_translated<NX>(a, (int64_t)st0 + fi, ox, oy, xf);
      hq = dp_lower_bound_time(S, xf, a.goal);
      const double cm = chosen != -1 ? chosen * S.dt : (L1 - 1) * S.dt;
      double nd[NX], t0[NX];
#pragma unroll
      for (int c = 0; c < NX; ++c) nd[c] = a.nodes[(int64_t)c * a.Ncap + i];
      load_translated<NX>(a, (int64_t)st0, ox, oy, t0);
      gq = a.gpar[i] + cm + a.cdf * dp_lower_bound_time(S, nd, t0);
    } else {
#pragma unroll
      for (int c = 0; c < NX; ++c) xf[c] = 0.0;
    }
    bool dupv = false;
    if (a.nv > 0 && !a.select && __ballot(fr)) {
      bool hit = false;
      // the 9 runs over the pair's lanes (one round from 16 lanes per pair)
      for (int k = k0; fr && k < 9 && !hit; k += LPP)
        hit = dp_grid_run_any<NX>(S, a.grid, a.vstart, a.vis, a.vstride, xf, k, a.r_dup, Tdup);
      dupv = ((__ballot(hit) >> shift) & segbits) != 0;
    }
    if (vp && k0 == 0) {
      a.verdict[q] = vd;
      a.firstk[q] = vd == DP_V_BOUNDS ? first_bad : (vd == DP_V_COLLIDE || vd == DP_V_CONSTRAINT ? first_col : -1);
      a.goal_out[q] = chosen;
      a.res[q] = fr && !a.select ? DP_R_UNDECIDED : DP_R_NONE;
      a.flags[q] = (int8_t)((dupv ? 1 : 0) | (chosen != -1 ? 2 : 0));
      if (fr) {
#pragma unroll
        for (int c = 0; c < NX; ++c) a.fin[(int64_t)c * a.pcap + q] = xf[c];
        a.gout[q] = gq;
        a.hout[q] = hq;
      }
      c_b += vd == DP_V_BOUNDS;
      c_c += vd == DP_V_COLLIDE;
      c_x += vd == DP_V_CONSTRAINT;
      c_f += fr;
      c_k += vd != DP_V_BOUNDS;
    }
  }
  if (k0 == 0) {
    if (c_x) atomicAdd(&s_cnt[0], (unsigned long long)c_x);
    if (c_b) atomicAdd(&s_cnt[1], (unsigned long long)c_b);
    if (c_c) atomicAdd(&s_cnt[2], (unsigned long long)c_c);
    if (c_f) atomicAdd(&s_cnt[3], (unsigned long long)c_f);
    if (c_k) atomicAdd(&s_cnt[4], (unsigned long long)c_k);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int slot[5] = {DPC_CONSTRAINT, DPC_BOUNDS, DPC_COLLIDE, DPC_FREE, DPC_CHECKED};
    for (int k = 0; k < 5; ++k)
      if (s_cnt[k]) atomicAdd((unsigned long long*)&a.counters[slot[k]], s_cnt[k]);
  }
}

// dbrrt's choice per node (dbrrt.cpp:627-660): among the node's FREE pairs,
// in pair order, the first with the smallest distance(last placed state,
// target) below DBL_MAX (the reference starts from
// numeric_limits<double>::max() and keeps a candidate only when d < min), or
// the first FREE one with choose_first.  One wave per node.  Writes the pair
// index (-1: none) and [last state, first state, distance, g, h].
template <int NX>
__global__ __launch_bounds__(64) void k_select(dp_check_args a, const int32_t* __restrict__ app, int32_t N,
                                               const double* __restrict__ targets, const double* __restrict__ cost,
                                               double cost_jump, int32_t choose_first, int32_t* __restrict__ sel_q,
                                               int32_t* __restrict__ sel_prim, double* __restrict__ sel) {
  const dp_dev_sys& S = *a.sys;
  const int32_t i = blockIdx.x;
  if (i >= N) return;
  const int lane = threadIdx.x;
  const int32_t tot = *a.total;
  const int32_t b = app[i] < tot ? app[i] : tot;
  const int32_t e = app[i + 1] < tot ? app[i + 1] : tot;
  double tg[NX];
#pragma unroll
  for (int c = 0; c < NX; ++c) tg[c] = targets[(int64_t)c * a.Ncap + i];
  const double ox = a.nodes[i], oy = a.nodes[(int64_t)a.Ncap + i];
  double best_d = DBL_MAX;
  int32_t best_q = -1;
  for (int32_t cb = b; cb < e; cb += 64) {
    const int32_t q = cb + lane;
    double d = DBL_MAX;
    bool ok = false;
    if (q < e && a.verdict[q] == DP_V_FREE) {
      const int32_t j = a.pprim[q];
      double x[NX];
      load_translated<NX>(a, (int64_t)a.off[j + 1] - 1, ox, oy, x);
      d = dp_distance(S, x, tg);
      ok = d < DBL_MAX;  // NaN and +inf never enter
    }
    if (choose_first) {
      const uint64_t m = __ballot(ok);
      if (m) {
        const int src = __ffsll((long long)m) - 1;
        best_d = __shfl(d, src);
        best_q = cb + src;
        break;
      }
      continue;
    }
    // first minimum of the chunk: (d, q) lexicographic over the lanes that qualify
    double md = ok ? d : DBL_MAX;
    int32_t mq = ok ? q : INT_MAX;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
      const double od = __shfl_xor(md, w);
      const int32_t oq = __shfl_xor(mq, w);
      if (od < md || (od == md && oq < mq)) {
        md = od;
        mq = oq;
      }
    }
    if (mq != INT_MAX && md < best_d) {  // strictly below every earlier chunk's minimum
      best_d = md;
      best_q = mq;
    }
  }
  if (lane == 0) {
    sel_q[i] = best_q;
    sel_prim[i] = best_q >= 0 ? a.pprim[best_q] : -1;
  }
  double* o = sel + (int64_t)i * (2 * NX + 3);
  if (best_q < 0) {
    if (lane < 2 * NX + 3) o[lane] = NAN;
    return;
  }
  if (lane == 0) {
    const int32_t j = a.pprim[best_q];
    double x[NX], x0[NX], nd[NX];
    load_translated<NX>(a, (int64_t)a.off[j + 1] - 1, ox, oy, x);
    load_translated<NX>(a, (int64_t)a.off[j], ox, oy, x0);
#pragma unroll
    for (int c = 0; c < NX; ++c) nd[c] = a.nodes[(int64_t)c * a.Ncap + i];
#pragma unroll
    for (int c = 0; c < NX; ++c) {
      o[c] = x[c];
      o[NX + c] = x0[c];
    }
    o[2 * NX] = best_d;
    o[2 * NX + 1] = a.gpar[i] + cost[j] + cost_jump * dp_lower_bound_time(S, nd, x0);  // dbrrt.cpp:645-648
    o[2 * NX + 2] = dp_lower_bound_time(S, x, a.goal);                                  // dbrrt.cpp:641-642
  }
}

}  // namespace

// ==========================================================================
// host-side pipeline
// ==========================================================================

template <typename F>
static int dispatch_nx(int nx, F&& f) {
  switch (nx) {
    case 3: return f(std::integral_constant<int, 3>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 5: return f(std::integral_constant<int, 5>{});
    case 13: return f(std::integral_constant<int, 13>{});
  }
  dp_set_error("unsupported nx");
  return DP_ERR_ARG;
}

// Applicability count only: the exact number of pairs of the uploaded nodes
// lands in counters[DPC_PAIRS] (buffers are sized from it).
// Exact rejection bounds of the applicability scan for radius r.
static app_bounds make_app_bounds(const dp_ctx* c, double r) {
  app_bounds B{INFINITY, INFINITY, INFINITY, -1.0};
  const dp_system_desc& d = c->desc;
  auto sq = [&](double w) {
    if (!(w > 0.0)) return (double)INFINITY;
    const double t = r / w;
    return t * t * (1.0 + 1e-9);
  };
  B.T_p = sq(d.dist_w[0]);
  if (d.system == DP_SYS_QUADROTOR) {
    B.T_v = sq(d.dist_w[2]);
    B.T_w = sq(d.dist_w[3]);
    // w1 * acos(|dot|) > r when |dot| < cos(r / w1); the 1e-9 margin covers
    // libm's cos and the few ulps of the device's acos series
    if (d.dist_w[1] > 0.0 && r / d.dist_w[1] < 1.5) B.c_q = std::cos(r / d.dist_w[1]) - 1e-9;
  }
  return B;
}

// Applicability masks and counts of the uploaded nodes (app cleared here).
template <int NX>
static void launch_app(dp_ctx* c, const app_bounds& B, double r_app, hipStream_t s);

// Applicability radius: alpha * delta as the float product of dbastar.cpp:585-586,
// or dbrrt's Expander radius (options_dbrrt.delta) for dp_expand_select.
static double app_radius(const dp_ctx* c, const dp_expand_params* p) {
  return c->select_mode ? c->select_radius : (double)(p->alpha * p->delta);
}

// Margins of the FP32 rejections (app_bounds32): relative 1e-4 on the
// compared quantities, absolute 1e-6 on top.
static app_bounds32 make_app_bounds32(const dp_ctx* c, double r) {
  const dp_system_desc& d = c->desc;
  app_bounds32 B{-1.0f, INFINITY, (float)d.dist_w[1], (float)d.dist_w[3], INFINITY};
  if (d.system == DP_SYS_QUADROTOR) {
    if (d.dist_w[1] > 0.0 && r / d.dist_w[1] < 1.5) B.c_q = (float)(std::cos(r / d.dist_w[1]) - 1e-5);
    if (d.dist_w[3] > 0.0) B.T_w = (float)((r / d.dist_w[3]) * (r / d.dist_w[3]) * (1.0 + 1e-4) + 1e-6);
    B.R = (float)(r * (1.0 + 1e-4) + 1e-6);
  } else {
    B.R = (float)(r * (1.0 + 1e-4) + d.dist_w[1] * 1e-6 + 1e-6);
  }
  return B;
}

// Primitive parts per node of the blocked applicability: about 4096 waves
// over the (node group, part) grid, at most one part per mask word.
static int32_t app_parts(const dp_ctx* c) {
  const int32_t groups = (c->N + kAppNodes - 1) / kAppNodes;
  const int32_t words = (c->P + 63) / 64;
  int32_t parts = (4096 + groups - 1) / groups;
  return parts < 1 ? 1 : (parts > words ? words : parts);
}
static bool app_per_node_order(const dp_ctx* c) { return c->d_order && c->order_rows > 1; }

template <int NX>
static void launch_app(dp_ctx* c, const app_bounds& B, double r_app, hipStream_t s) {
  const int32_t N = c->N;
  dp_fill_i32(c->d_app, N, 0, s);
  const double* keys = c->forward ? c->d_start : c->d_last;
  if (app_per_node_order(c)) {  // a pair order per node: one wave per node part
    k_app_count<NX><<<N * kAppParts, 64, 0, s>>>(c->d_sys, B, c->d_nodes, c->N_cap, N, keys, c->P, r_app, c->d_order,
                                                 c->order_rows, c->d_app, c->d_app_mask);
    return;
  }
  const int32_t groups = (N + kAppNodes - 1) / kAppNodes;
  const int32_t parts = app_parts(c);
  const float* keys32 = c->forward ? c->d_start32 : c->d_last32;
  k_app_blk<NX><<<groups * parts, 64, 0, s>>>(c->d_sys, make_app_bounds32(c, r_app), c->d_nodes, c->N_cap, N, keys,
                                              keys32, c->P, r_app, c->d_order, parts, c->d_app, c->d_app_mask);
}

int dp_launch_count(dp_ctx* c, const dp_expand_params* p, hipStream_t s) {
  const double r_app = app_radius(c, p);
  const int32_t N = c->N;
  const app_bounds B = make_app_bounds(c, r_app);
  return dispatch_nx(c->desc.nx, [&](auto NXc) -> int {
    constexpr int NX = decltype(NXc)::value;
    launch_app<NX>(c, B, r_app, s);
    dp_scan_small(c->d_app, N, c->d_total, c->pair_cap, c->d_counters, DPC_PAIRS, s);
    DP_HIP_RET(hipGetLastError());
    return DP_OK;
  });
}

// Stage events are skipped while a multi-instance graph is being captured.
#define DP_EV(ev_, s_) \
  do { \
    if (!c->no_events) DP_HIP_RET(hipEventRecord(ev_, s_)); \
  } while (0)

int dp_launch_expand(dp_ctx* c, const dp_expand_params* p, hipStream_t s) {
  const int nx = c->desc.nx;
  const int32_t N = c->N;
  const double r_app = app_radius(c, p);                                 // dbastar.cpp:585-586 (float product)
  const double r_dup = (1. - (double)p->alpha) * (double)p->delta;      // dbastar.cpp:852
  const double goal_bound = p->delta_factor_goal * (double)p->delta;    // dbastar.cpp:827

  if (!c->select_mode && (c->grid_dirty || c->grid_radius != r_dup)) {
    int rc = dp_launch_visited_rebuild(c, r_dup, s);
    if (rc) return rc;
  }

  DP_EV(c->ev[0], s);
  dp_zero_counters(c->d_counters, s);
  const app_bounds B = make_app_bounds(c, r_app);
  int rc = dispatch_nx(nx, [&](auto NXc) -> int {
    constexpr int NX = decltype(NXc)::value;
    launch_app<NX>(c, B, r_app, s);
    dp_scan_small(c->d_app, N, c->d_total, c->pair_cap, c->d_counters, DPC_PAIRS, s);
    if (c->d_order)
      k_app_fill<true><<<N, 64, 0, s>>>(N, c->P, c->d_app, c->d_app_mask, c->d_order, c->order_rows, c->d_pnode,
                                        c->d_pprim, c->pair_cap);
    else
      k_app_fill<false><<<N, 64, 0, s>>>(N, c->P, c->d_app, c->d_app_mask, nullptr, 1, c->d_pnode, c->d_pprim,
                                         c->pair_cap);
    return DP_OK;
  });
  if (rc) return rc;
  DP_EV(c->ev[4], s);
  if (p->use_surrogate) {
    if (!c->mlp.w || (p->use_surrogate == 1 && !c->mlp.ready)) {
      dp_set_error("use_surrogate needs a trained and calibrated surrogate");
      return DP_ERR_STATE;
    }
    if (c->desc.system == DP_SYS_QUADROTOR) {
      dp_set_error("the surrogate is defined for planar systems");
      return DP_ERR_ARG;
    }
    rc = dp_launch_surrogate(c, p->use_surrogate == 1, s);
    if (rc) return rc;
  }
  DP_EV(c->ev[1], s);

  dp_check_args a{};
  a.sys = c->d_sys;
  a.nodes = c->d_nodes;
  a.Ncap = c->N_cap;
  a.gpar = c->d_gpar;
  a.off = c->d_off;
  a.st = c->d_st;
  a.S = c->S;
  a.geo = c->d_geo;
  a.nbox = c->nbox;
  a.sweep = c->d_sweep;
  a.srange = c->d_srange;
  a.P = c->P;
  a.goalidx = c->d_goalidx;
  a.num_goal = p->num_check_goal;
  a.check_goal = p->check_intermediate_goal && p->num_check_goal > 0;
  a.obs = c->d_obs;
  a.nobs = c->nb;
  a.node_obs = c->d_node_obs;
  a.node_nobs = c->d_node_nobs;
  a.vis = c->d_vis_sorted;
  a.vstride = c->nv_cap;
  a.nv = c->nv;
  a.vstart = c->d_vis_start;
  a.grid = c->grid;
  a.pnode = c->d_pnode;
  a.pprim = c->d_pprim;
  a.total = c->d_total;
  a.verdict = c->d_verdict;
  a.firstk = c->d_firstk;
  a.goal_out = c->d_goal;
  a.res = c->d_res;
  a.flags = c->d_flags;
  a.fin = c->d_final;
  a.pcap = c->pair_cap;
  a.gout = c->d_g;
  a.hout = c->d_h;
  for (int k = 0; k < DP_MAX_NX; ++k) a.goal[k] = k < nx ? p->goal[k] : 0.0;
  a.goal_bound = goal_bound;
  a.r_dup = r_dup;
  a.cdf = p->cost_delta_factor;
  a.surrogate = p->use_surrogate ? 1 : 0;
  a.counters = c->d_counters;
  a.backward = c->forward ? 0 : 1;
  a.last = c->d_last;
  a.ncon = c->select_mode ? 0 : c->ncon;
  a.con = c->d_con;
  a.con_delta = (double)p->delta;
  a.select = c->select_mode ? 1 : 0;
  if (c->select_mode) a.nv = 0;

  if (c->desc.system == DP_SYS_QUADROTOR) {
    rc = dp_launch_quad(c, p, goal_bound, r_dup, s);
    if (rc) return rc;
    DP_EV(c->ev[2], s);
    rc = dp_launch_resolve(c, p, r_dup, s);
    if (rc) return rc;
    DP_EV(c->ev[3], s);
    c->timing_valid = !c->no_events && !c->defer_walk;
    return DP_OK;
  }
  int lpp = c->tiles.lanes_per_pair > 0 ? c->tiles.lanes_per_pair
                                       : (c->maxlen <= 16 ? 16 : (c->maxlen <= 32 ? 32 : 64));
  // LDS: the obstacles plus one candidate list per pair group; narrow shapes
  // widen until that fits a workgroup
  auto lds_of = [&](int l) {
    return sizeof(double) * 4 * c->nb + sizeof(int16_t) * (size_t)kWaves * (64 / l) * (c->nb > 0 ? c->nb : 1);
  };
  while (lpp < 64 && lds_of(lpp) > (size_t)DP_LDS_MAX - sizeof(dp_dev_sys)) lpp *= 2;
  const int seg = 64 / lpp;
  const size_t lds = lds_of(lpp);
  // The pair count is on the device; size the grid from the bound N*P.
  const int64_t groups = ((int64_t)N * c->P + seg - 1) / seg;
  const int blocks = dp_grid_blocks(groups, kWaves, c->tiles.check_blocks);
  if (N > 0) {
    const double* rb = c->forward ? c->reach_box : c->reach_box_bwd;
    k_node_obs<<<dp_grid_blocks(N, kWaves, 1024), 256, 0, s>>>(c->d_nodes, c->N_cap, N, c->d_obs, c->nb, rb[0], rb[1],
                                                              rb[2], rb[3], c->d_node_obs, c->d_node_nobs);
    DP_HIP_RET(hipGetLastError());
  }
  if ((rc = dp_record_event(c, 5, s))) return rc;
  rc = dispatch_nx(nx, [&](auto NXc) -> int {
    constexpr int NX = decltype(NXc)::value;
    if constexpr (NX == 13) {
      return DP_ERR_ARG;
    } else {
      if (lpp == 2) k_check_planar<NX, 2><<<blocks, 256, lds, s>>>(a);
      else if (lpp == 4) k_check_planar<NX, 4><<<blocks, 256, lds, s>>>(a);
      else if (lpp == 8) k_check_planar<NX, 8><<<blocks, 256, lds, s>>>(a);
      else if (lpp == 16) k_check_planar<NX, 16><<<blocks, 256, lds, s>>>(a);
      else if (lpp == 32) k_check_planar<NX, 32><<<blocks, 256, lds, s>>>(a);
      else k_check_planar<NX, 64><<<blocks, 256, lds, s>>>(a);
      return DP_OK;
    }
  });
  if (rc) return rc;
  DP_HIP_RET(hipGetLastError());
  DP_EV(c->ev[6], s);
  DP_EV(c->ev[2], s);

  if (c->select_mode) {
    rc = dispatch_nx(nx, [&](auto NXc) -> int {
      constexpr int NX = decltype(NXc)::value;
      if constexpr (NX == 13) {
        return DP_ERR_ARG;
      } else {
        k_select<NX><<<N, 64, 0, s>>>(a, c->d_app, N, c->d_targets, c->d_cost, c->select_cost_jump,
                                      c->select_first, c->d_sel_q, c->d_sel_q + c->sel_cap, c->d_sel);
        return DP_OK;
      }
    });
    if (rc) return rc;
    DP_HIP_RET(hipGetLastError());
    DP_EV(c->ev[7], s);
    DP_EV(c->ev[3], s);
    c->timing_valid = !c->no_events && !c->defer_walk;
    return DP_OK;
  }
  rc = dp_launch_resolve(c, p, r_dup, s);
  if (rc) return rc;
  DP_EV(c->ev[3], s);
  c->timing_valid = !c->no_events && !c->defer_walk;
  return DP_OK;
}
