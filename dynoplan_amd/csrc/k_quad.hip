// k_quad.hip -- quadrotor_v0 expansion: rollout of the primitive's actions
// from the translated start state (transform_primitive with the position /
// velocity offset, dbastar.hpp:166-215), bounds on every state, sphere robot
// against boxes and spheres with the first colliding state
// (check_lazy_trajectory, dbastar.cpp:391-490), then check_goal, g/h and the
// duplicate test on the FP64 states.
//
// Kernels:
//   k_quad32   one wave per (node, chunk of 64 pairs): FP32 rollout relative
//              to the node, obstacles of the node's reach box staged in LDS
//              as FP32 offsets from the node, per-state clearance.  A state
//              whose clearance (or bound margin) is within near_eps is
//              "near"; a verdict that a near state could change is deferred
//              to the FP64 pass.  Quaternions are renormalised in FP64.
//   k_quad64   one lane per deferred pair and per FREE pair: the FP64 rollout
//              in the same operation order as the CPU oracle (bit-identical
//              states), full FP64 re-check for deferred pairs, then goal
//              check, g, h and the visited-set duplicate test.
//   k_quad_window  one wave per node, between the two: FREE pairs past the
//              node's FP64 window are deferred (no FP64 rollout unless the
//              node's cap does not fall before them).
// With fp32_rollout = 0, k_quad64 evaluates every pair (exact FP64 mode).
#include <hip/hip_runtime.h>
#include <limits.h>

#include "dp_internal.h"

// flag bits of d_flags (bit0 dup-vs-visited, bit1 goal-forced are set later)
#define QF_NEAR 4
#define QF_NEED64 8

// What the FP32 pass reads of primitive j, in one 128-byte record (one
// cache line) instead of a dozen scattered arrays: start velocity (FP64: the
// pair's start velocity is (float)(v0 + node velocity)), start position as
// FP32, the rate-bound / near states, the state range, the reach, and the
// offset box of the rollout up to the rate-bound state and over all states
// (k_quad_rot).
struct alignas(16) quad_rec {
  double v0[3];
  float p0[3];
  int32_t rbad, rnear, s0, L1;
  float reach;
  float qbox[12];
  float dbox[6];  // position offset D min / max over every state 0..L1 - 1
};
static_assert(sizeof(quad_rec) == 128, "one cache line per primitive");

struct dp_quad_args {
  const dp_dev_sys* sys;
  const double* nodes;
  int32_t Ncap, N;
  const double* gpar;
  const int32_t* off;
  const double* st;       // [13][S]
  int32_t S;
  const double* start;    // [13][P] start state of each primitive (the pairs read it, not st)
  const double* act;      // [A][4]
  const float* act32;     // [A][4]
  int32_t A;
  int32_t P;
  const int32_t* goalidx;
  int32_t num_goal, check_goal;
  const double* obs;      // [nb][6] then [ns][4]
  int32_t nb, ns;
  double prim_reach;      // max over primitives of max_k |p_k - p_0|_inf
  double tmax;            // longest primitive duration
  const float* preach;    // [P] per primitive: max_k |p_k - p_0|_inf + (L - 1) dt, the reach at unit speed offset
  const double* vis;
  int64_t vstride, nv;
  const int32_t* vstart;
  dp_grid_geom grid;
  const int32_t* app;     // pair offsets per node [N+1]
  const int32_t* chunk_off;  // [N+1] wave chunks per node, scanned
  const int32_t* chunk_node;  // [chunks] node of each wave chunk
  const int2* qperm;      // [pairs] (pair, primitive) of each node's pairs, those the primitive's rate bounds reject first
  const int32_t* nshort;  // [N] how many of them
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
  int32_t exact;
  int32_t redo;           // k_quad64: only the pairs deferred past their FP64 window
  int32_t* q64;           // FP32 mode: the pairs of the FP64 pass (counters[DPC_Q64] of them)
  // per primitive state s: attitude / rate trajectory from the stored start
  // (k_quad_rot), shared by every pair of the primitive
  const double* rot64;    // [7][S] qx qy qz qw wx wy wz
  const double* dv64;     // [S][4] velocity increment of the step into state s (x y z, pad)
  const float* dv32;      // [S][4] the same in FP32
  const int32_t* rbad64;  // [P] first state whose attitude / rate leaves the bounds (INT_MAX: none)
  const int32_t* rbad32;  // [P] the same in FP32
  const int32_t* rnear32; // [P] first FP32 state within the re-check margin of an attitude / rate bound
  const quad_rec* qrec;   // [P] the FP32 pass's view of each primitive
  long long* counters;
};

namespace {

constexpr int kQWaves = 4;
// obstacles listed per lane for the FP64 re-check (int16 in LDS)
constexpr int kQ64Cand = 128;

__device__ __forceinline__ bool collides64(double r, const double* obs, int nb, int ns, const double* s) {
  for (int o = 0; o < nb; ++o)
    if (dp_sphere_box(s[0], s[1], s[2], r, obs + 6 * o)) return true;
  const double* sp = obs + 6 * nb;
  for (int o = 0; o < ns; ++o)
    if (dp_sphere_sphere(s[0], s[1], s[2], r, sp + 4 * o)) return true;
  return false;
}

// ---- FP32 pass -------------------------------------------------------------------
struct quad_f32 {
  float g4, g, armF, t2tF, Jx, Jy, Jz, dt, h;
  float cosc[6], sincc[6];
};

// FP32 attitude / rate part of the step (quaternion renormalised in FP64)
// and the FP32 velocity increment dv (the state's velocity becomes v + dv).
// It reads only s[3..6], s[10..12] and the rotors, so it is evaluated once
// per primitive and shared by every pair (k_quad_rot).
__device__ __forceinline__ void quad_rot_step32(const quad_f32& K, float* s, const float* u, float* dv) {
  const float qx = s[3], qy = s[4], qz = s[5], qw = s[6];
  const float wx = s[10], wy = s[11], wz = s[12];
  const float f = K.g4 * (((u[0] + u[1]) + u[2]) + u[3]);
  const float tx = K.armF * (((u[2] + u[3]) - u[0]) - u[1]);
  const float ty = K.armF * (((u[1] + u[2]) - u[0]) - u[3]);
  const float tz = K.t2tF * (((u[1] + u[3]) - u[0]) - u[2]);
  const float rzx = 2.0f * (qx * qz + qw * qy);
  const float rzy = 2.0f * (qy * qz - qw * qx);
  const float rzz = 1.0f - 2.0f * (qx * qx + qy * qy);
  const float jwx = K.Jx * wx, jwy = K.Jy * wy, jwz = K.Jz * wz;
  const float dwx = (tx - (wy * jwz - wz * jwy)) / K.Jx;
  const float dwy = (ty - (wz * jwx - wx * jwz)) / K.Jy;
  const float dwz = (tz - (wx * jwy - wy * jwx)) / K.Jz;
  dv[0] = K.dt * (f * rzx);
  dv[1] = K.dt * (f * rzy);
  dv[2] = K.dt * (f * rzz - K.g);
  s[10] = wx + K.dt * dwx;
  s[11] = wy + K.dt * dwy;
  s[12] = wz + K.dt * dwz;
  const float z = (K.h * K.h) * ((wx * wx + wy * wy) + wz * wz);
  float pc = K.cosc[5], ps = K.sincc[5];
#pragma unroll
  for (int n = 4; n >= 0; --n) {
    pc = pc * z + K.cosc[n];
    ps = ps * z + K.sincc[n];
  }
  const float kk = K.h * ps;
  const float ex = kk * wx, ey = kk * wy, ez = kk * wz, ew = pc;
  const double nw = ((qw * ew - qx * ex) - qy * ey) - qz * ez;
  const double nx = ((qw * ex + qx * ew) + qy * ez) - qz * ey;
  const double ny = ((qw * ey - qx * ez) + qy * ew) + qz * ex;
  const double nz = ((qw * ez + qx * ey) - qy * ex) + qz * ew;
  // quaternion renormalised in FP64
  const double inv = 1.0 / sqrt(((nx * nx + ny * ny) + nz * nz) + nw * nw);
  s[3] = (float)(nx * inv);
  s[4] = (float)(ny * inv);
  s[5] = (float)(nz * inv);
  s[6] = (float)(nw * inv);
}

__device__ __forceinline__ quad_f32 make_quad_f32(const dp_dev_sys& S) {
  quad_f32 K;
  K.g = (float)S.quad[1];
  K.g4 = (float)(S.quad[1] * 0.25);
  K.armF = (float)(S.quad[2] * (S.quad[0] * (S.quad[1] * 0.25)));
  K.t2tF = (float)(S.quad[3] * (S.quad[0] * (S.quad[1] * 0.25)));
  K.Jx = (float)S.quad[4];
  K.Jy = (float)S.quad[5];
  K.Jz = (float)S.quad[6];
  K.dt = (float)S.dt;
  K.h = (float)(0.5 * S.dt);
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    K.cosc[n] = (float)S.cos_c[n];
    K.sincc[n] = (float)S.sinc_c[n];
  }
  return K;
}

// Per primitive, once per upload: the attitude / rate rollout from the
// stored start and the velocity increments of its steps, in FP64 (the FP64
// step's operations, dp_quad_rot_step64) and in FP32 (quad_rot_step32), plus
// the first state whose attitude or rate leaves the bounds (FP64 exact;
// FP32 with the re-check margins, also the first state within them).  The
// position / velocity offset of a pair does not enter these components, so
// every pair of the primitive shares them.  One thread per primitive.
__global__ void k_quad_rot(const dp_dev_sys* __restrict__ Sp, int32_t P, const int32_t* __restrict__ off,
                           const double* __restrict__ st, int32_t Sn, const double* __restrict__ act,
                           const float* __restrict__ act32, double* __restrict__ rot64, double* __restrict__ dv64,
                           float* __restrict__ dv32, int32_t* __restrict__ rbad64, int32_t* __restrict__ rbad32,
                           int32_t* __restrict__ rnear32, const float* __restrict__ preach,
                           quad_rec* __restrict__ qrec) {
  const dp_dev_sys& S = *Sp;
  const quad_f32 K = make_quad_f32(S);
  const float eps_q = 100.0f * S.near_eps, eps_w = 100.0f * S.near_eps;
  for (int32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < P; j += gridDim.x * blockDim.x) {
    const int32_t s0 = off[j];
    const int32_t L1 = off[j + 1] - s0;
    const int32_t a0 = s0 - j;
    double x[13];
    float y[13];
#pragma unroll
    for (int c = 0; c < 13; ++c) {
      x[c] = st[(int64_t)c * Sn + s0];
      y[c] = (float)x[c];
    }
    int32_t b64 = INT_MAX, b32 = INT_MAX, n32 = INT_MAX;
    for (int k = 0; k < L1; ++k) {
      if (k > 0) {
        double dv[3];
        const double4 uv = reinterpret_cast<const double4*>(act)[a0 + k - 1];
        const double u[4] = {uv.x, uv.y, uv.z, uv.w};
        double o[13];
        dp_quad_rot_step64(S, x, u, o, dv);
#pragma unroll
        for (int c = 3; c < 7; ++c) x[c] = o[c];
#pragma unroll
        for (int c = 10; c < 13; ++c) x[c] = o[c];
        float dvf[3];
        const float4 uf = reinterpret_cast<const float4*>(act32)[a0 + k - 1];
        const float u32[4] = {uf.x, uf.y, uf.z, uf.w};
        quad_rot_step32(K, y, u32, dvf);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          dv64[(int64_t)(s0 + k) * 4 + c] = c < 3 ? dv[c] : 0.0;
          dv32[(int64_t)(s0 + k) * 4 + c] = c < 3 ? dvf[c] : 0.0f;
        }
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          dv64[(int64_t)s0 * 4 + c] = 0.0;
          dv32[(int64_t)s0 * 4 + c] = 0.0f;
        }
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) rot64[(int64_t)c * Sn + s0 + k] = x[3 + c];
#pragma unroll
      for (int c = 0; c < 3; ++c) rot64[(int64_t)(4 + c) * Sn + s0 + k] = x[10 + c];
      bool bad64 = false, bad32 = false, near32 = false;
#pragma unroll
      for (int c = 3; c < 13; ++c) {
        if (c >= 7 && c < 10) continue;  // velocities depend on the pair
        bad64 = bad64 || !(x[c] >= S.lo[c] && x[c] <= S.hi[c]);
        const float lo_ = (float)S.lo[c], hi_ = (float)S.hi[c];
        const float e = c < 7 ? eps_q : eps_w;
        bad32 = bad32 || !(y[c] >= lo_ && y[c] <= hi_);
        near32 = near32 || fabsf(y[c] - lo_) <= e || fabsf(y[c] - hi_) <= e;
      }
      if (bad64 && b64 == INT_MAX) b64 = k;
      if (bad32 && b32 == INT_MAX) b32 = k;
      if (near32 && n32 == INT_MAX) n32 = k;
    }
    rbad64[j] = b64;
    rbad32[j] = b32;
    rnear32[j] = n32;
    quad_rec rec{};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rec.v0[c] = st[(int64_t)(7 + c) * Sn + s0];
      rec.p0[c] = (float)st[(int64_t)c * Sn + s0];
    }
    rec.rbad = b32;
    rec.rnear = n32;
    rec.s0 = s0;
    rec.L1 = L1;
    rec.reach = preach[j];
    // Offsets of the FP32 rollout over states 0..min(b32, L1 - 1), without
    // the pair's start velocity v0: v_k = v0 + C_k (C_k the sum of the first
    // k FP32 increments) and p_k = p0 + k dt v0 + D_k (D_k = dt times the sum
    // of C_0..C_{k-1}).  Summed in FP64 from the FP32 increments, widened by
    // 1e-3 (m, m/s): the FP32 loop's own rounding stays below 1e-4 over 50
    // steps at these magnitudes.
    {
      const int32_t kend = b32 < L1 ? b32 : L1 - 1;
      double C[3] = {0.0, 0.0, 0.0}, D[3] = {0.0, 0.0, 0.0};
      double cmin[3] = {0.0, 0.0, 0.0}, cmax[3] = {0.0, 0.0, 0.0}, dmin[3] = {0.0, 0.0, 0.0},
             dmax[3] = {0.0, 0.0, 0.0}, fmn[3] = {0.0, 0.0, 0.0}, fmx[3] = {0.0, 0.0, 0.0};
      const double dt = S.dt;
      for (int k = 1; k < L1; ++k) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          D[c] += dt * C[c];
          C[c] += (double)dv32[(int64_t)(s0 + k) * 4 + c];
          fmn[c] = fmin(fmn[c], D[c]);
          fmx[c] = fmax(fmx[c], D[c]);
          if (k <= kend) {
            cmin[c] = fmin(cmin[c], C[c]);
            cmax[c] = fmax(cmax[c], C[c]);
            dmin[c] = fmin(dmin[c], D[c]);
            dmax[c] = fmax(dmax[c], D[c]);
          }
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        rec.dbox[c] = (float)(fmn[c] - 1e-3);
        rec.dbox[3 + c] = (float)(fmx[c] + 1e-3);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        rec.qbox[c] = (float)(cmin[c] - 1e-3);
        rec.qbox[3 + c] = (float)(cmax[c] + 1e-3);
        rec.qbox[6 + c] = (float)(dmin[c] - 1e-3);
        rec.qbox[9 + c] = (float)(dmax[c] + 1e-3);
      }
    }
    qrec[j] = rec;
  }
}

// FP32 start state of pair (node, primitive j), relative to the node:
// stored p0 and v0 plus the node velocity (the position / velocity offset).
__device__ __forceinline__ void quad32_start(const dp_quad_args& a, int32_t j, const double* nvel, float* s) {
#pragma unroll
  for (int c = 0; c < 13; ++c) {
    const double v = a.start[(int64_t)c * a.P + j];
    s[c] = (c >= 7 && c <= 9) ? (float)(v + nvel[c - 7]) : (float)v;
  }
}

__global__ __launch_bounds__(256) void k_quad32(dp_quad_args a) {
  extern __shared__ float qsm[];
  const dp_dev_sys& S = *a.sys;
  // the wave index is uniform: in a scalar register, every per-chunk value
  // derived from it (chunk, node, its offsets and state) is a scalar load
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
  const int per_wave = 6 * a.nb + 4 * a.ns;
  float* cb = qsm + wave * per_wave;  // compacted boxes (6 floats) then spheres (4 floats)
  // the workgroup's counts, added to the batch counters once per workgroup
  // (once per wave, thousands of waves queued on three addresses at the end)
  __shared__ unsigned long long s_cnt[3];
  if (threadIdx.x < 3) s_cnt[threadIdx.x] = 0ull;
  __syncthreads();
  const quad_f32 K = make_quad_f32(S);
  const float r = (float)S.radius;
  // margins under which a verdict is re-checked in FP64: near_eps for
  // positions (clearance and position bounds), 10x for velocities and 100x
  // for angular rates, whose FP32 drift is larger (measured on primitives
  // of 10-50 steps: 1e-6 m, 9e-6 m/s, 2.4e-5 rad/s at most)
  const float eps = S.near_eps, eps_v = 10.0f * S.near_eps, eps_w = 100.0f * S.near_eps;
  const float r_lo2 = (r - eps) * (r - eps), r_hi2 = (r + eps) * (r + eps), r2 = r * r;
  const int32_t total_chunks = a.chunk_off[a.N];
  long long c_b = 0, c_c = 0, c_k = 0;

  for (int32_t w = blockIdx.x * kQWaves + wave; w < total_chunks; w += gridDim.x * kQWaves) {
    const int32_t node = a.chunk_node[w];
    const int32_t tot = *a.total;
    const int32_t pbeg = a.app[node] < tot ? a.app[node] : tot;
    const int32_t pend = a.app[node + 1] < tot ? a.app[node + 1] : tot;
    // the node's chunks: first those of pairs whose primitive leaves the
    // rate bounds (BOUNDS whatever the translation: no collision tests), then
    // the others
    const int32_t ns = a.nshort[node];
    const int32_t cshort = (ns + 63) / 64;
    const int32_t cw = w - a.chunk_off[node];
    const bool short_chunk = cw < cshort;
    const int32_t idx = short_chunk ? pbeg + cw * 64 + lane : pbeg + ns + (cw - cshort) * 64 + lane;
    const bool vp = short_chunk ? idx < pbeg + ns : idx < pend;
    const int2 qj = vp ? a.qperm[idx] : make_int2(0, 0);
    const int32_t q = qj.x;
    double np[3], nvel[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      np[c] = a.nodes[(int64_t)c * a.Ncap + node];
      nvel[c] = a.nodes[(int64_t)(7 + c) * a.Ncap + node];
    }
    // Reach box of this chunk's rollouts (positions relative to the node):
    // the largest displacement of its primitives plus their duration times
    // the node's speed, with a 0.05 m margin.  Only obstacles meeting it are
    // staged; a state leaving it defers the pair to the FP64 re-check, which
    // tests every obstacle that meets the pair's own path.
    const double vmax = fmax(fabs(nvel[0]), fmax(fabs(nvel[1]), fabs(nvel[2])));
    double reach = 0.0;
    if (!short_chunk) {
      float rl = 0.0f;
      if (vp) {
        const quad_rec& rj = a.qrec[qj.y];
        rl = rj.reach + (float)((rj.L1 - 1) * a.sys->dt * vmax);
      }
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) rl = fmaxf(rl, __shfl_xor(rl, d, 64));
      reach = (double)rl + 0.05;
    }
    const float freach = (float)reach;
    // obstacles are staged out to the robot radius plus 2 near_eps beyond the
    // box: an FP32 state inside it is then more than r + near_eps from every
    // obstacle left out, on some axis, and so is its FP64 counterpart
    const double stage_pad = S.radius + 2.0 * S.near_eps;
    // stage candidate obstacles relative to the node, ballot-compacted
    int nbc = 0, nsc = 0;
    for (int o0 = 0; o0 < (short_chunk ? 0 : a.nb); o0 += 64) {
      const int o = o0 + lane;
      bool keep = false;
      float rel[6];
      if (o < a.nb) {
        const double* b = a.obs + 6 * o;
        keep = true;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double d = b[c] - np[c];
          keep = keep && fabs(d) <= reach + b[3 + c] + stage_pad;
          rel[c] = (float)d;
          rel[3 + c] = (float)b[3 + c];
        }
      }
      const uint64_t m = __ballot(keep);
      if (keep) {
        float* dst = cb + 6 * (nbc + __popcll(m & ((1ull << lane) - 1ull)));
#pragma unroll
        for (int c = 0; c < 6; ++c) dst[c] = rel[c];
      }
      nbc += __popcll(m);
    }
    float* cs = cb + 6 * a.nb;
    for (int o0 = 0; o0 < (short_chunk ? 0 : a.ns); o0 += 64) {
      const int o = o0 + lane;
      bool keep = false;
      float rel[4];
      if (o < a.ns) {
        const double* b = a.obs + 6 * a.nb + 4 * o;
        keep = true;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double d = b[c] - np[c];
          keep = keep && fabs(d) <= reach + b[3] + stage_pad;
          rel[c] = (float)d;
        }
        rel[3] = (float)b[3];
      }
      const uint64_t m = __ballot(keep);
      if (keep) {
        float* dst = cs + 4 * (nsc + __popcll(m & ((1ull << lane) - 1ull)));
#pragma unroll
        for (int c = 0; c < 4; ++c) dst[c] = rel[c];
      }
      nsc += __popcll(m);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    if (vp) {
      const quad_rec& rj = a.qrec[qj.y];
      const int32_t s0 = rj.s0;
      const int32_t L1 = rj.L1;
      // position (relative to the node) and velocity; the attitude / rate
      // part of every step is the primitive's (k_quad_rot)
      float s[6];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        s[c] = rj.p0[c];
        s[3 + c] = (float)(rj.v0[c] + nvel[c]);
      }
      const int32_t rbad = rj.rbad, rnear = rj.rnear;
      float plo[3], phi[3], vlo[3], vhi[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        plo[c] = (float)(S.lo[c] - np[c]);
        phi[c] = (float)(S.hi[c] - np[c]);
        vlo[c] = (float)S.lo[7 + c];
        vhi[c] = (float)S.hi[7 + c];
      }
      int first_bad = INT_MAX, first_col = INT_MAX;
      bool near_bound_any = false, near_col_before = false, near = false;
      // A rate-bound pair whose position and velocity stay, over states
      // 0..rbad, more than the margins inside their bounds (from the
      // primitive's offset box and the pair's start) ends BOUNDS at rbad,
      // near exactly when an attitude / rate state before it is: the
      // rollout below would find the same.
      bool fast = false;
      if (short_chunk) {
        const float* qb = rj.qbox;
        const float R = (float)rbad * K.dt;
        fast = true;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float v0 = s[3 + c];
          const float vl = v0 + qb[c], vh = v0 + qb[3 + c];
          const float rv = R * v0;
          const float pl = s[c] + qb[6 + c] + fminf(rv, 0.0f), ph = s[c] + qb[9 + c] + fmaxf(rv, 0.0f);
          fast = fast && vl > vlo[c] + 2.0f * eps_v && vh < vhi[c] - 2.0f * eps_v;
          fast = fast && pl > plo[c] + 2.0f * eps && ph < phi[c] - 2.0f * eps;
        }
        if (fast) {
          first_bad = rbad;
          near_bound_any = rnear <= rbad;
        }
      }
      // The velocity increments of the next four steps are in flight while
      // a state is checked, in four registers that the loop below (unrolled
      // by four) names statically: the increment of step k sits in slot
      // (k - 1) % 4 and is replaced by that of step k + 4 right after use,
      // so each step waits for a load issued four steps earlier.  Indices
      // are clamped to the primitive's last state, so every load is valid
      // and unconditional (a conditional load would make the wait for it
      // exact only at the loop's end).
      // Pairs decided from their primitive's record (the large majority)
      // all read the table's first entry: one cache line for the wave.
      const float4* dvp = reinterpret_cast<const float4*>(a.dv32) + (fast ? 0 : s0);
      const int32_t Lr = fast ? 0 : L1;
      const int32_t Lm = fast ? 0 : L1 - 1;
      // (issued in the loop's own order, slot 3 first, so that the wait
      // for each slot is the same on entry and around the loop)
      float4 dq[4];
      dq[3] = dvp[4 < Lm ? 4 : Lm];
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        __builtin_amdgcn_sched_barrier(0);  // keep the issue order
        dq[u] = dvp[1 + u < Lm ? 1 + u : Lm];
      }
      __builtin_amdgcn_sched_barrier(0);
      // Long pairs: the staged obstacles that can come within r + 2 near_eps
      // of this pair's own states (its position box: the primitive's offset
      // box plus k dt v0 for k up to L1 - 1, widened by 1e-3), one bit each;
      // an obstacle left out is farther than that from every state, so it
      // neither collides nor comes near.  Without room for the bits, all.
      uint64_t bmask = ~0ull, smask = ~0ull;
      const bool masked = !short_chunk && nbc + nsc <= 64;
      if (masked) {
        const float* qb = rj.qbox;
        const float R = (float)(L1 - 1) * K.dt;
        const float pad = r + 2.0f * eps + 1e-3f;
        float lo[3], hi[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float rv = R * s[3 + c];
          lo[c] = s[c] + qb[6 + c] + fminf(rv, 0.0f) - pad;
          hi[c] = s[c] + qb[9 + c] + fmaxf(rv, 0.0f) + pad;
        }
        bmask = 0ull;
        for (int t = 0; t < nbc; ++t) {
          const float* b = cb + 6 * t;
          bool hit = true;
#pragma unroll
          for (int c = 0; c < 3; ++c) hit = hit && b[c] - b[3 + c] <= hi[c] && b[c] + b[3 + c] >= lo[c];
          if (hit) bmask |= 1ull << t;
        }
        smask = 0ull;
        for (int t = 0; t < nsc; ++t) {
          const float* b = cs + 4 * t;
          bool hit = true;
#pragma unroll
          for (int c = 0; c < 3; ++c) hit = hit && b[c] - b[3] <= hi[c] && b[c] + b[3] >= lo[c];
          if (hit) smask |= 1ull << t;
        }
      }
      bool stop = false;
      for (int k0 = 0; k0 < Lr && !stop; k0 += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + u;
        const int slot = (u + 3) & 3;
        const bool live = !stop && k < Lr;
        if (live && k > 0) {
          const float dv[3] = {dq[slot].x, dq[slot].y, dq[slot].z};
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            s[c] = s[c] + K.dt * s[3 + c];
            s[3 + c] = s[3 + c] + dv[c];
          }
        }
        dq[slot] = dvp[k + 4 < Lm ? k + 4 : Lm];
        if (!live) continue;
        // bounds with margin (attitude and rate: the primitive's first
        // offending / near state)
        bool bad = k >= rbad, nb_ = k >= rnear;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          bad = bad || !(s[c] >= plo[c] && s[c] <= phi[c]);
          nb_ = nb_ || fabsf(s[c] - plo[c]) <= eps || fabsf(s[c] - phi[c]) <= eps;
          nb_ = nb_ || (!short_chunk && fabsf(s[c]) > freach);  // left the staged reach box
          bad = bad || !(s[3 + c] >= vlo[c] && s[3 + c] <= vhi[c]);
          nb_ = nb_ || fabsf(s[3 + c] - vlo[c]) <= eps_v || fabsf(s[3 + c] - vhi[c]) <= eps_v;
        }
        near_bound_any = near_bound_any || nb_;
        if (bad) {
          first_bad = k;
          stop = true;  // later states cannot change a bounds verdict
          continue;
        }
        if (first_col == INT_MAX && !short_chunk) {
          bool col = false, nc = false;
          auto box_test = [&](int t) {
            const float* b = cb + 6 * t;
            float dx = fabsf(s[0] - b[0]) - b[3];
            float dy = fabsf(s[1] - b[1]) - b[4];
            float dz = fabsf(s[2] - b[2]) - b[5];
            dx = dx > 0.f ? dx : 0.f;
            dy = dy > 0.f ? dy : 0.f;
            dz = dz > 0.f ? dz : 0.f;
            const float d2 = (dx * dx + dy * dy) + dz * dz;
            col = col || d2 <= r2;
            nc = nc || (d2 >= r_lo2 && d2 <= r_hi2);
          };
          auto sphere_test = [&](int t) {
            const float* b = cs + 4 * t;
            const float dx = s[0] - b[0], dy = s[1] - b[1], dz = s[2] - b[2];
            const float d2 = (dx * dx + dy * dy) + dz * dz;
            const float rr = r + b[3];
            col = col || d2 <= rr * rr;
            nc = nc || (d2 >= (rr - eps) * (rr - eps) && d2 <= (rr + eps) * (rr + eps));
          };
          if (masked) {
            for (uint64_t m = bmask; m; m &= m - 1ull) box_test(__ffsll((long long)m) - 1);
            for (uint64_t m = smask; m; m &= m - 1ull) sphere_test(__ffsll((long long)m) - 1);
          } else {
            for (int t = 0; t < nbc; ++t) box_test(t);
            for (int t = 0; t < nsc; ++t) sphere_test(t);
          }
          if (nc) near_col_before = true;
          if (col) first_col = k;
        }
      }
      }
      int8_t vd;
      int32_t fk;
      if (first_bad != INT_MAX) {
        vd = DP_V_BOUNDS;
        fk = first_bad;
        near = near_bound_any;
      } else if (first_col != INT_MAX) {
        vd = DP_V_COLLIDE;
        fk = first_col;
        near = near_bound_any || near_col_before;
      } else {
        vd = DP_V_FREE;
        fk = -1;
        near = near_bound_any || near_col_before;
      }
      if (near) {
        a.flags[q] = QF_NEAR;
      } else if (vd == DP_V_FREE) {
        a.flags[q] = QF_NEED64;
      } else {
        a.flags[q] = 0;
        a.verdict[q] = vd;
        a.firstk[q] = fk;
        a.goal_out[q] = -1;
        a.res[q] = DP_R_NONE;
        c_b += vd == DP_V_BOUNDS;
        c_c += vd == DP_V_COLLIDE;
        c_k += vd == DP_V_COLLIDE;
      }
    }
    // the staging area is rewritten by the next chunk
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (c_b) atomicAdd(&s_cnt[0], (unsigned long long)c_b);
  if (c_c) atomicAdd(&s_cnt[1], (unsigned long long)c_c);
  if (c_k) atomicAdd(&s_cnt[2], (unsigned long long)c_k);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_cnt[0]) atomicAdd((unsigned long long*)&a.counters[DPC_BOUNDS], s_cnt[0]);
    if (s_cnt[1]) atomicAdd((unsigned long long*)&a.counters[DPC_COLLIDE], s_cnt[1]);
    if (s_cnt[2]) atomicAdd((unsigned long long*)&a.counters[DPC_CHECKED], s_cnt[2]);
  }
}

// ---- FP64 pass -------------------------------------------------------------------
// Pairs of the FP64 pass: in FP32 mode the compact list k_quad_window built
// (every lane of a wave busy), else every pair of the batch.
__device__ __forceinline__ int64_t quad64_count(const dp_quad_args& a) {
  return a.exact || a.redo ? (int64_t)*a.total : (int64_t)a.counters[DPC_Q64];
}
__device__ __forceinline__ int64_t quad64_pair(const dp_quad_args& a, int64_t t) {
  return a.exact || a.redo ? t : (int64_t)a.q64[t];
}

// Start state of pair (node i, primitive j) in FP64: stored state plus the
// node's position / velocity (the offset dimensions).
__device__ __forceinline__ void quad64_start(const dp_dev_sys& S, const dp_quad_args& a, int32_t i, int32_t j,
                                             double* s) {
#pragma unroll
  for (int c = 0; c < 13; ++c) {
    const double v = a.start[(int64_t)c * a.P + j];
    s[c] = dp_is_offset_dim(S, c) ? v + a.nodes[(int64_t)c * a.Ncap + i] : v;
  }
}

// State k of the pair from state k-1: translation with the primitive's
// velocity increment (dp_quad_trans_step64); s[3..6], s[10..12] untouched.
__device__ __forceinline__ void quad64_advance(const dp_quad_args& a, int32_t s0, int k, double* s) {
  const double4 d = reinterpret_cast<const double4*>(a.dv64)[s0 + k];
  const double dv[3] = {d.x, d.y, d.z};
  double o[13];
  dp_quad_trans_step64(a.sys->dt, s, dv, o);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    s[c] = o[c];
    s[7 + c] = o[7 + c];
  }
}

// Attitude and rate of primitive state st into s[3..6], s[10..12].
__device__ __forceinline__ void quad64_rot(const dp_quad_args& a, int64_t st, double* s) {
#pragma unroll
  for (int c = 0; c < 4; ++c) s[3 + c] = a.rot64[(int64_t)c * a.S + st];
#pragma unroll
  for (int c = 0; c < 3; ++c) s[10 + c] = a.rot64[(int64_t)(4 + c) * a.S + st];
}

// Bounds of the position and velocity components (is_state_valid on them),
// against lo[6] / hi[6] (position then velocity) held in registers; every
// comparison is evaluated (no short circuit, so no branch per bound).
__device__ __forceinline__ bool trans_valid64(const double* lo, const double* hi, const double* s) {
  bool ok = true;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    ok = ok & (s[c] >= lo[c]) & (s[c] <= hi[c]);
    ok = ok & (s[7 + c] >= lo[3 + c]) & (s[7 + c] <= hi[3 + c]);
  }
  return ok;
}

// FP64 pass, one lane per pair.  A pair that needs the full re-check is
// rolled out twice: first for the bounds over every state (the reference
// rejects on bounds before it tests collisions, dbastar.cpp:440-446) and the
// box of its positions; if in bounds, the obstacles meeting that box
// (inflated by the robot radius and 1e-9 m) are listed in the lane's LDS
// slice, and the second rollout tests each state against them only.  Both
// rollouts form the same states.  Other pairs are rolled out once, for the
// goal checks, the final state, g and h.  The velocity increments of the
// next four steps are in flight during each step: with about one wave per
// SIMD the pass is bound by load latency.
__global__ __launch_bounds__(256) void k_quad64(dp_quad_args a) {
  // LDS: the obstacles (boxes [nb][6] then spheres [ns][4]), then the
  // per-lane candidate lists [blockDim.x][kQ64Cand]
  extern __shared__ double q64_lds[];
  const dp_dev_sys& S = *a.sys;
  const int64_t n64 = quad64_count(a);
  if ((int64_t)blockIdx.x * blockDim.x >= n64) return;  // uniform per block
  const int nobs_d = 6 * a.nb + 4 * a.ns;
  double* obs = q64_lds;
  for (int t = threadIdx.x; t < nobs_d; t += blockDim.x) obs[t] = a.obs[t];
  // the workgroup's counts, added to the batch counters once per workgroup
  __shared__ unsigned long long s_cnt[5];
  if (threadIdx.x < 5) s_cnt[threadIdx.x] = 0ull;
  __syncthreads();
  // read once into registers: the rollout loops store to global memory,
  // after which a field of *a.sys would be read again on every step
  const double dt = S.dt, rad = S.radius, w0 = S.w[0];
  const double Tdup = dp_pos_bound(S, a.r_dup);
  double blo[6], bhi[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    blo[c] = S.lo[c];
    bhi[c] = S.hi[c];
    blo[3 + c] = S.lo[7 + c];
    bhi[3 + c] = S.hi[7 + c];
  }
  int16_t* cand = reinterpret_cast<int16_t*>(q64_lds + nobs_d) + (int64_t)threadIdx.x * kQ64Cand;
  long long c_b = 0, c_c = 0, c_f = 0, c_k = 0, c_r = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n64; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = quad64_pair(a, t);
    int8_t fl = a.exact ? (int8_t)QF_NEAR : a.flags[q];
    if (a.redo) fl = (fl & DP_F_DEFER) ? (int8_t)QF_NEED64 : (int8_t)0;
    if (!(fl & (QF_NEAR | QF_NEED64))) continue;
    const bool full = (fl & QF_NEAR) != 0;
    if (full && !a.exact) ++c_r;
    const int32_t i = a.pnode[q], j = a.pprim[q];
    const int32_t s0 = a.off[j];
    const int32_t L1 = a.off[j + 1] - s0;
    // s[0..2] position, s[7..9] velocity; attitude and rate come from the
    // primitive's trajectory (k_quad_rot) when a full state is needed
    double s[13];
    int first_bad = INT_MAX, first_col = INT_MAX;
    const int32_t rbad = a.rbad64[j];
    int ncand = 0;  // obstacles listed for the second rollout (-1: more than kQ64Cand, test all)
    // The velocity increments of the next four steps are in flight while a
    // state is processed, in four statically named slots of loops unrolled
    // by four (step k's increment in slot (k - 1) % 4, replaced by step
    // k + 4's right after use); indices clamp to the last state, so every
    // load is valid and unconditional, and the slots are first loaded in
    // the loops' own order (3, 0, 1, 2): each step then waits for a load
    // issued four steps earlier.
    const double4* dvp = reinterpret_cast<const double4*>(a.dv64) + s0;
    const int32_t Lm = L1 - 1;
    auto ld = [&](int idx) {
      const double4 t = dvp[idx < Lm ? idx : Lm];
      return make_double3(t.x, t.y, t.z);
    };
    auto advance = [&](const double3& d) {
      const double dv[3] = {d.x, d.y, d.z};
      double o[13];
      dp_quad_trans_step64(dt, s, dv, o);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        s[c] = o[c];
        s[7 + c] = o[7 + c];
      }
    };
    double3 dq[4];
    if (full) {
      // first rollout: bounds of every state and the box of the positions
      quad64_start(S, a, i, j, s);
      double bl[3] = {s[0], s[1], s[2]}, bh[3] = {s[0], s[1], s[2]};
      dq[3] = ld(4);
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        __builtin_amdgcn_sched_barrier(0);  // keep the issue order
        dq[u] = ld(1 + u);
      }
      __builtin_amdgcn_sched_barrier(0);
      bool stop = false;
      for (int k0 = 0; k0 < L1 && !stop; k0 += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int k = k0 + u;
          const int slot = (u + 3) & 3;
          const bool live = !stop && k < L1;
          if (live && k > 0) advance(dq[slot]);
          dq[slot] = ld(k + 4);
          if (!live) continue;
          if (k >= rbad || !trans_valid64(blo, bhi, s)) {
            first_bad = k;
            stop = true;
            continue;
          }
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            bl[c] = fmin(bl[c], s[c]);
            bh[c] = fmax(bh[c], s[c]);
          }
        }
      }
      if (first_bad == INT_MAX) {
        // a sphere of radius r at p meets a box only if p lies in the box
        // grown by r on every axis; pad 1e-9 m against rounding
        const double rr = S.radius + 1e-9;
        for (int o = 0; o < a.nb + a.ns && ncand >= 0; ++o) {
          const bool box = o < a.nb;
          const double* b = box ? obs + 6 * o : obs + 6 * a.nb + 4 * (o - a.nb);
          bool meet = true;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const double h = (box ? b[3 + c] : b[3]) + rr;
            meet = meet && b[c] - h <= bh[c] && b[c] + h >= bl[c];
          }
          if (meet) {
            if (ncand < kQ64Cand) cand[ncand++] = (int16_t)o;
            else ncand = -1;
          }
        }
      }
    }
    // The node and the first translated state only enter g: its lower-bound
    // term is formed here, so neither stays live through the rollout.
    double lb0;
    {
      double node[13];
#pragma unroll
      for (int c = 0; c < 13; ++c) node[c] = a.nodes[(int64_t)c * a.Ncap + i];
      quad64_start(S, a, i, j, s);
      lb0 = dp_lower_bound_time(S, node, s);
    }
    int chosen = -1, nn = 0;
    const int nng = a.check_goal ? a.num_goal : 0;
    // the goal indices of the pair's primitive, loaded before the rollout
    // (at most 8; a load per step in the loop would serialise on its latency)
    int32_t gi[8];
#pragma unroll
    for (int g = 0; g < 8; ++g) gi[g] = g < nng ? a.goalidx[(int64_t)g * a.P + j] : -1;
    int32_t gnext = nng > 0 ? gi[0] : -1;
    dq[3] = ld(4);
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      __builtin_amdgcn_sched_barrier(0);  // keep the issue order
      dq[u] = ld(1 + u);
    }
    __builtin_amdgcn_sched_barrier(0);
    const int32_t L2 = first_bad == INT_MAX ? L1 : 0;
    for (int k0 = 0; k0 < L2; k0 += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + u;
      const int slot = (u + 3) & 3;
      const bool live = k < L2;
      if (live && k > 0) advance(dq[slot]);
      dq[slot] = ld(k + 4);
      if (!live) continue;
      if (full && first_col == INT_MAX) {
        bool col;
        if (ncand < 0) {
          col = collides64(rad, obs, a.nb, a.ns, s);
        } else {
          col = false;
          for (int e = 0; e < ncand && !col; ++e) {
            const int o = cand[e];
            col = o < a.nb ? dp_sphere_box(s[0], s[1], s[2], rad, obs + 6 * o)
                           : dp_sphere_sphere(s[0], s[1], s[2], rad, obs + 6 * a.nb + 4 * (o - a.nb));
          }
        }
        if (col) first_col = k;
      }
      // check_goal at the float-rule indices, in nn order (dbastar.cpp:492-512)
      while (chosen == -1 && nn < nng && gnext == k) {
        // dp_within's own first exit, the position term above the bound,
        // taken before the attitude is loaded: no load in the common case
        if (!(w0 * dp_l2_3(s, a.goal) > a.goal_bound)) {
          quad64_rot(a, s0 + k, s);
          if (dp_within(S, s, a.goal, a.goal_bound)) {
            chosen = k;
            // the final state is the goal state: stored now, read back below
#pragma unroll
            for (int c = 0; c < 13; ++c) a.fin[(int64_t)c * a.pcap + q] = s[c];
          }
        }
        ++nn;
#pragma unroll
        for (int g = 1; g < 8; ++g)
          if (nn == g) gnext = gi[g];
      }
    }
    }
    if (first_bad == INT_MAX) quad64_rot(a, s0 + L1 - 1, s);  // the full final state
    int8_t vd = DP_V_FREE;
    int32_t fk = -1;
    if (first_bad != INT_MAX) {
      vd = DP_V_BOUNDS;
      fk = first_bad;
    } else if (first_col != INT_MAX) {
      vd = DP_V_COLLIDE;
      fk = first_col;
    }
    a.verdict[q] = vd;
    a.firstk[q] = fk;
    if (!a.redo) {  // deferred pairs were counted by k_quad_window
      c_b += vd == DP_V_BOUNDS;
      c_c += vd == DP_V_COLLIDE;
      c_k += vd != DP_V_BOUNDS;
    }
    if (vd != DP_V_FREE) {
      a.goal_out[q] = -1;
      a.res[q] = DP_R_NONE;
      a.flags[q] = 0;
      continue;
    }
    double x[13];
    if (chosen == -1) {
#pragma unroll
      for (int c = 0; c < 13; ++c) x[c] = s[c];
    } else {
#pragma unroll
      for (int c = 0; c < 13; ++c) x[c] = a.fin[(int64_t)c * a.pcap + q];
    }
    const double hq = dp_lower_bound_time(S, x, a.goal);
    const double cm = chosen != -1 ? chosen * S.dt : (L1 - 1) * S.dt;
    const double gq = a.gpar[i] + cm + a.cdf * lb0;
#pragma unroll
    for (int c = 0; c < 13; ++c) a.fin[(int64_t)c * a.pcap + q] = x[c];
    a.gout[q] = gq;
    a.hout[q] = hq;
    a.goal_out[q] = chosen;
    a.res[q] = DP_R_UNDECIDED;
    // duplicate of the visited set: the 9 runs of the final state's 27-cell
    // block, with the state still in registers
    bool dupv = false;
    if (a.nv > 0)
      for (int k = 0; k < 9 && !dupv; ++k)
        dupv = dp_grid_run_any<13>(S, a.grid, a.vstart, a.vis, a.vstride, x, k, a.r_dup, Tdup);
    a.flags[q] = (int8_t)((chosen != -1 ? 2 : 0) | (dupv ? 1 : 0));
    c_f += !a.redo;
  }
  if (c_b) atomicAdd(&s_cnt[0], (unsigned long long)c_b);
  if (c_c) atomicAdd(&s_cnt[1], (unsigned long long)c_c);
  if (c_f) atomicAdd(&s_cnt[2], (unsigned long long)c_f);
  if (c_k) atomicAdd(&s_cnt[3], (unsigned long long)c_k);
  if (c_r) atomicAdd(&s_cnt[4], (unsigned long long)c_r);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int slot[5] = {DPC_BOUNDS, DPC_COLLIDE, DPC_FREE, DPC_CHECKED, DPC_RECHECK};
    for (int k = 0; k < 5; ++k)
      if (s_cnt[k]) atomicAdd((unsigned long long*)&a.counters[slot[k]], s_cnt[k]);
  }
}

// Per node: its pairs reordered for the FP32 pass, those whose primitive
// leaves the attitude / rate bounds within its length first (their verdict
// is BOUNDS whatever the translation, so their rollouts are short and need
// no collision test), then the rest, each group in canonical order (the
// results are written nearly in order); and the node's wave chunks (64
// pairs, one group per chunk).  One wave per node.
__global__ __launch_bounds__(64) void k_quad_split(int32_t N, const int32_t* __restrict__ app,
                                                   const int32_t* __restrict__ total, const int32_t* __restrict__ pprim,
                                                   const int32_t* __restrict__ off, const int32_t* __restrict__ rbad32,
                                                   int2* __restrict__ qperm, int32_t* __restrict__ nshort,
                                                   int32_t* __restrict__ chunks) {
  const int32_t node = blockIdx.x;
  if (node >= N) return;
  const int lane = threadIdx.x;
  const uint64_t lt = (1ull << lane) - 1ull;
  const int32_t tot = *total;
  const int32_t b = app[node] < tot ? app[node] : tot;
  const int32_t e = app[node + 1] < tot ? app[node + 1] : tot;
  int32_t ns = 0;
  for (int32_t cb = b; cb < e; cb += 64) {
    const int32_t p = cb + lane;
    bool sh = false;
    if (p < e) {
      const int32_t j = pprim[p];
      sh = rbad32[j] < off[j + 1] - off[j];
    }
    ns += __popcll(__ballot(sh));
  }
  int32_t ws = b, wl = b + ns;
  for (int32_t cb = b; cb < e; cb += 64) {
    const int32_t p = cb + lane;
    bool sh = false;
    int32_t j = 0;
    if (p < e) {
      j = pprim[p];
      sh = rbad32[j] < off[j + 1] - off[j];
    }
    const uint64_t ms = __ballot(sh), ml = __ballot(p < e && !sh);
    if (p < e) qperm[sh ? ws + __popcll(ms & lt) : wl + __popcll(ml & lt)] = make_int2(p, j);
    ws += __popcll(ms);
    wl += __popcll(ml);
  }
  if (lane == 0) {
    nshort[node] = ns;
    chunks[node] = (ns + 63) / 64 + (e - b - ns + 63) / 64;
  }
}

// Node of every wave chunk (after the scan of the chunk counts).
__global__ void k_quad_chunk_nodes(int32_t N, const int32_t* __restrict__ chunk_off, int32_t* __restrict__ chunk_node) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
    for (int32_t w = chunk_off[i]; w < chunk_off[i + 1]; ++w) chunk_node[w] = i;
}

// FP64 window of each node: the pairs the FP32 pass found FREE are ranked in
// canonical order (near-contact pairs, re-checked in FP64 whatever their
// place, are not counted), so a pair's rank never exceeds its rank among
// the node's FREE pairs; a FREE pair at rank >= window is deferred: no FP64 rollout, no
// final state, goal index, g or h unless its node's cap does not fall before
// it (the resolution then reports a window overflow and the batch is
// resolved again with every deferred pair rolled out).  The reference
// never computes a pair after its node's cap (dbastar.cpp:916-918).  A
// deferred pair's verdict (FREE) is exact from the FP32 pass.
// The window is `window` FREE pairs, doubled wshift[0] times (at most 2^10)
// unless `fixed`: the resolution raises wshift[0] when a node overflows, so a
// context's later batches stay on the device; wshift[1] records the growth
// this batch used.
__global__ __launch_bounds__(256) void k_quad_window(int32_t N, const int32_t* __restrict__ app,
                                                     const int32_t* __restrict__ total, int8_t* __restrict__ verdict,
                                                     int32_t* __restrict__ firstk, int32_t* __restrict__ goal_out,
                                                     int8_t* __restrict__ res, int8_t* __restrict__ flags,
                                                     int32_t window, int32_t fixed, int32_t* __restrict__ wshift,
                                                     int32_t* __restrict__ q64, long long* __restrict__ counters) {
  // four nodes per workgroup, one per wave: their places in the compact list
  // are reserved with one atomic per workgroup (one per node, all issued
  // together on one address, queue behind each other)
  __shared__ int32_t s_need[4];
  __shared__ int32_t s_base;
  const int wv = threadIdx.x / 64;
  const int32_t node = blockIdx.x * 4 + wv;
  const int lane = threadIdx.x % 64;
  if (!fixed) {
    const int32_t sh = wshift[0] < 10 ? wshift[0] : 10;
    const int64_t w = (int64_t)window << sh;
    window = w < INT32_MAX ? (int32_t)w : INT32_MAX;
    if (node == 0 && lane == 0) wshift[1] = sh;
  }
  const uint64_t lt = (1ull << lane) - 1ull;
  const int32_t tot = *total;
  const int32_t b = node < N ? (app[node] < tot ? app[node] : tot) : 0;
  const int32_t e = node < N ? (app[node + 1] < tot ? app[node + 1] : tot) : 0;
  // first pass: deferrals, and the node's count of pairs for the FP64 pass
  int32_t rank = 0, need_n = 0;
  long long nd = 0;
  // each chunk's flags are loaded while the previous chunk is processed
  // (indices clamped to the node's last pair: unconditional loads, and the
  // one past the last chunk goes unused); a lane reads and writes only the
  // pairs p = b + lane (mod 64), so the second pass sees its own stores
  const int32_t el = e > b ? e - 1 : b;
  int8_t fl_n = e > b ? flags[b + lane < el ? b + lane : el] : (int8_t)0;
  for (int32_t cb = b; cb < e; cb += 64) {
    const int32_t p = cb + lane;
    const int8_t fl_c = fl_n;
    fl_n = flags[cb + 64 + lane < el ? cb + 64 + lane : el];
    const int8_t fl = p < e ? fl_c : (int8_t)0;
    const uint64_t m = __ballot((fl & QF_NEED64) != 0);
    const bool defer = (fl & QF_NEED64) && rank + __popcll(m & lt) >= window;
    if (defer) {
      flags[p] = (int8_t)DP_F_DEFER;
      verdict[p] = DP_V_FREE;
      firstk[p] = -1;
      goal_out[p] = -1;
      res[p] = DP_R_UNDECIDED;
      ++nd;
    }
    need_n += __popcll(__ballot((fl & (QF_NEAR | QF_NEED64)) != 0 && !defer));
    rank += __popcll(m);
  }
  if (lane == 0) s_need[wv] = need_n;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int32_t sum = s_need[0] + s_need[1] + s_need[2] + s_need[3];
    s_base = sum ? (int32_t)atomicAdd((unsigned long long*)&counters[DPC_Q64], (unsigned long long)sum) : 0;
  }
  __syncthreads();
  if (need_n > 0) {
    int32_t base = s_base;
    for (int k = 0; k < wv; ++k) base += s_need[k];
    // second pass: the pairs in canonical order (deferred ones now carry DP_F_DEFER)
    fl_n = flags[b + lane < el ? b + lane : el];
    for (int32_t cb = b; cb < e; cb += 64) {
      const int32_t p = cb + lane;
      const int8_t fl_c = fl_n;
      fl_n = flags[cb + 64 + lane < el ? cb + 64 + lane : el];
      const int8_t fl = p < e ? fl_c : (int8_t)0;
      const bool need = (fl & (QF_NEAR | QF_NEED64)) != 0;
      const uint64_t nm = __ballot(need);
      if (need) q64[base + __popcll(nm & lt)] = p;
      base += __popcll(nm);
    }
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) nd += __shfl_xor(nd, d, 64);
  if (lane == 0 && nd) {
    atomicAdd((unsigned long long*)&counters[DPC_DEFERRED], (unsigned long long)nd);
    atomicAdd((unsigned long long*)&counters[DPC_FREE], (unsigned long long)nd);
    atomicAdd((unsigned long long*)&counters[DPC_CHECKED], (unsigned long long)nd);
  }
}

// Every state of the listed pairs of the last batch: FP32 as k_quad32 forms
// them (the same step function, position made absolute in FP64), or FP64 as
// k_quad64 forms them; out [n][maxlen][13], rows past a primitive's length
// untouched.  One thread per pair; error reports only.
__global__ void k_quad64_trace(dp_quad_args a, const int32_t* __restrict__ pairs, int32_t n, int32_t maxlen,
                               double* __restrict__ out) {
  const dp_dev_sys& S = *a.sys;
  for (int32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int32_t q = pairs[t];
    const int32_t node = a.pnode[q], j = a.pprim[q];
    const int32_t s0 = a.off[j];
    const int32_t L1 = a.off[j + 1] - s0;
    double s[13];
    quad64_start(S, a, node, j, s);
    for (int k = 0; k < L1 && k < maxlen; ++k) {
      if (k > 0) quad64_advance(a, s0, k, s);
      quad64_rot(a, s0 + k, s);
      double* w = out + ((int64_t)t * maxlen + k) * 13;
      for (int c = 0; c < 13; ++c) w[c] = s[c];
    }
  }
}

// FP32: the same step functions as k_quad_rot and k_quad32 (attitude and
// rate rolled out per pair here, bit-identical to the shared trajectory).
__global__ void k_quad32_trace(dp_quad_args a, const int32_t* __restrict__ pairs, int32_t n, int32_t maxlen,
                               double* __restrict__ out) {
  const dp_dev_sys& S = *a.sys;
  const quad_f32 K = make_quad_f32(S);
  for (int32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int32_t q = pairs[t];
    const int32_t node = a.pnode[q], j = a.pprim[q];
    const int32_t s0 = a.off[j];
    const int32_t L1 = a.off[j + 1] - s0;
    const int32_t a0 = s0 - j;
    double np[3], nvel[3];
    for (int c = 0; c < 3; ++c) {
      np[c] = a.nodes[(int64_t)c * a.Ncap + node];
      nvel[c] = a.nodes[(int64_t)(7 + c) * a.Ncap + node];
    }
    float s[13];
    quad32_start(a, j, nvel, s);
    for (int k = 0; k < L1 && k < maxlen; ++k) {
      if (k > 0) {
        const float4 uv = reinterpret_cast<const float4*>(a.act32)[a0 + k - 1];
        const float u[4] = {uv.x, uv.y, uv.z, uv.w};
        float dv[3];
        quad_rot_step32(K, s, u, dv);
        for (int c = 0; c < 3; ++c) {
          s[c] = s[c] + K.dt * s[7 + c];
          s[7 + c] = s[7 + c] + dv[c];
        }
      }
      double* o = out + ((int64_t)t * maxlen + k) * 13;
      for (int c = 0; c < 13; ++c) o[c] = c < 3 ? np[c] + (double)s[c] : (double)s[c];
    }
  }
}

dp_quad_args make_quad_args(dp_ctx* c, const dp_expand_params* p, double goal_bound, double r_dup) {
  dp_quad_args a{};
  a.sys = c->d_sys;
  a.nodes = c->d_nodes;
  a.Ncap = c->N_cap;
  a.N = c->N;
  a.gpar = c->d_gpar;
  a.off = c->d_off;
  a.st = c->d_st;
  a.S = c->S;
  a.start = c->d_start;
  a.act = c->d_act;
  a.act32 = c->d_act32;
  a.A = c->S - c->P;
  a.P = c->P;
  a.goalidx = c->d_goalidx;
  a.num_goal = p->num_check_goal;
  a.check_goal = p->check_intermediate_goal && p->num_check_goal > 0;
  a.obs = c->d_obs;
  a.nb = c->nb;
  a.ns = c->ns;
  a.prim_reach = c->prim_reach;
  a.preach = c->d_preach;
  a.tmax = (c->maxlen - 1) * c->desc.dt;
  a.vis = c->d_vis_sorted;
  a.vstride = c->nv_cap;
  a.nv = c->nv;
  a.vstart = c->d_vis_start;
  a.grid = c->grid;
  a.app = c->d_app;
  a.chunk_off = c->d_chunk;
  a.qperm = reinterpret_cast<const int2*>(c->d_qperm);
  a.chunk_node = c->d_chunk_node;
  a.nshort = c->d_nshort;
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
  for (int k = 0; k < DP_MAX_NX; ++k) a.goal[k] = k < 13 ? p->goal[k] : 0.0;
  a.goal_bound = goal_bound;
  a.r_dup = r_dup;
  a.cdf = p->cost_delta_factor;
  a.exact = p->fp32_rollout ? 0 : 1;
  a.redo = 0;
  a.q64 = c->d_q64;
  a.rot64 = c->d_rot64;
  a.dv64 = c->d_dv64;
  a.dv32 = c->d_dv32;
  a.rbad64 = c->d_rbad64;
  a.rbad32 = c->d_rbad32;
  a.rnear32 = c->d_rnear32;
  a.qrec = reinterpret_cast<const quad_rec*>(c->d_qrec);
  a.counters = c->d_counters;
  return a;
}

}  // namespace

int dp_launch_quad_rot(dp_ctx* c, hipStream_t s) {
  if (c->P <= 0) return DP_OK;
  k_quad_rot<<<dp_grid_blocks(c->P, 64, 4096), 64, 0, s>>>(c->d_sys, c->P, c->d_off, c->d_st, c->S, c->d_act,
                                                           c->d_act32, c->d_rot64, c->d_dv64, c->d_dv32, c->d_rbad64,
                                                           c->d_rbad32, c->d_rnear32, c->d_preach,
                                                           reinterpret_cast<quad_rec*>(c->d_qrec));
  DP_HIP_RET(hipGetLastError());
  return DP_OK;
}

// LDS of a k_quad64 workgroup of t64 lanes: the obstacles, the candidate lists.
static size_t quad64_lds(const dp_ctx* c, int t64) {
  return sizeof(double) * (size_t)(6 * c->nb + 4 * c->ns) + sizeof(int16_t) * (size_t)t64 * kQ64Cand;
}

int32_t dp_fp64_window(const dp_ctx* c, int32_t limit) {
  if (c->fp64_window > 0) return c->fp64_window;
  return limit >= 0 ? DP_FP64_WINDOW(limit) : INT32_MAX;
}

int dp_launch_quad_deferred(dp_ctx* c, hipStream_t s) {
  const dp_expand_params* p = &c->round_params;
  dp_quad_args a = make_quad_args(c, p, p->delta_factor_goal * (double)p->delta, c->round_r_dup);
  a.redo = 1;  // over every pair of the batch, the deferred ones rolled out
  const int t64 = c->tiles.quad64_threads;
  k_quad64<<<dp_grid_blocks(c->pair_cap, t64, c->tiles.quad64_blocks), t64, quad64_lds(c, t64), s>>>(a);
  DP_HIP_RET(hipGetLastError());
  return DP_OK;
}

extern "C" int dp_debug_rollout_states(dp_ctx* c, const int32_t* pairs, int32_t n, int32_t bits, double* out) {
  if (!c || (n > 0 && (!pairs || !out)) || n < 0 || (bits != 32 && bits != 64)) {
    dp_set_error("invalid argument");
    return DP_ERR_ARG;
  }
  if (c->desc.system != DP_SYS_QUADROTOR || c->N <= 0) {
    dp_set_error("no quadrotor batch expanded");
    return DP_ERR_STATE;
  }
  if (n == 0) return DP_OK;
  DP_HIP_RET(hipSetDevice(c->device));
  DP_HIP_RET(hipDeviceSynchronize());
  int32_t* d_pairs = nullptr;
  double* d_out = nullptr;
  const size_t out_bytes = sizeof(double) * (size_t)n * c->maxlen * 13;
  DP_HIP_RET(hipMalloc(&d_pairs, sizeof(int32_t) * n));
  hipError_t e = hipMalloc(&d_out, out_bytes);
  if (e == hipSuccess) e = hipMemcpy(d_pairs, pairs, sizeof(int32_t) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const dp_expand_params p = c->round_params;
    dp_quad_args a = make_quad_args(c, &p, 0.0, 0.0);
    if (bits == 32) k_quad32_trace<<<dp_grid_blocks(n, 256, 1024), 256, 0, c->stream>>>(a, d_pairs, n, c->maxlen, d_out);
    else k_quad64_trace<<<dp_grid_blocks(n, 256, 1024), 256, 0, c->stream>>>(a, d_pairs, n, c->maxlen, d_out);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(out, d_out, out_bytes, hipMemcpyDeviceToHost);
  (void)hipFree(d_pairs);
  (void)hipFree(d_out);
  DP_HIP_RET(e);
  return DP_OK;
}

int dp_launch_quad(dp_ctx* c, const dp_expand_params* p, double goal_bound, double r_dup, hipStream_t s) {
  dp_quad_args a = make_quad_args(c, p, goal_bound, r_dup);
  if (p->fp32_rollout) {
    const size_t lds = sizeof(float) * kQWaves * (6 * (size_t)c->nb + 4 * (size_t)c->ns + 1);
    if (lds > 64 * 1024) {
      dp_set_error("too many obstacles for the FP32 quadrotor staging (max 64 KiB LDS per workgroup)");
      return DP_ERR_ARG;
    }
    k_quad_split<<<c->N, 64, 0, s>>>(c->N, c->d_app, c->d_total, c->d_pprim, c->d_off, c->d_rbad32,
                                     reinterpret_cast<int2*>(c->d_qperm),
                                     c->d_nshort, c->d_chunk);
    dp_scan_small(c->d_chunk, c->N, nullptr, INT64_MAX, c->d_counters, -1, s);
    k_quad_chunk_nodes<<<dp_grid_blocks(c->N, 256, 1024), 256, 0, s>>>(c->N, c->d_chunk, c->d_chunk_node);
    const int64_t max_chunks = (int64_t)c->N * ((c->P + 63) / 64);
    int rc = dp_record_event(c, 5, s);
    if (rc) return rc;
    k_quad32<<<dp_grid_blocks(max_chunks, kQWaves, c->tiles.quad32_blocks), 256, lds, s>>>(a);
    if ((rc = dp_record_event(c, 6, s))) return rc;
    const int32_t window = dp_fp64_window(c, p->limit_branching_factor);
    const int32_t fixed = c->fp64_window > 0 || p->limit_branching_factor < 0 ? 1 : 0;
    if (c->N > 0)
      k_quad_window<<<(c->N + 3) / 4, 256, 0, s>>>(c->N, c->d_app, c->d_total, c->d_verdict, c->d_firstk, c->d_goal, c->d_res,
                                        c->d_flags, window, fixed, c->d_wshift, c->d_q64, c->d_counters);
  } else {
    int rc = dp_record_event(c, 5, s);
    if (!rc) rc = dp_record_event(c, 6, s);
    if (rc) return rc;
  }
  const int t64 = c->tiles.quad64_threads;
  k_quad64<<<dp_grid_blocks(c->pair_cap, t64, c->tiles.quad64_blocks), t64, quad64_lds(c, t64), s>>>(a);
  DP_HIP_RET(hipGetLastError());
  return DP_OK;
}
