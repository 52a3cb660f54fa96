// k_expand.hip -- gfx950 kernels of the batched db-A* expansion.
//
// Pipeline of one dp_expand (all on one stream, no host round trip):
//   k_app<count>      one wave per node: applicability query nearestR(c, alpha*delta)
//                     over primitive start states (Expander::expand_lazy,
//                     dbastar.hpp:270-312), brute force, ballot-counted
//   k_scan            exclusive scan -> canonical pair offsets per node
//   k_app<fill>       same scan, ballot-compacted (node, primitive) pair list
//   k_check_planar    LPP lanes per pair, one lane per primitive state:
//                     translate (transform_primitive), is_state_valid on every
//                     state, swept-box cull of LDS-staged obstacles, SAT
//                     per state, first offending index by ballot
//                     (check_lazy_trajectory, dbastar.cpp:391-490), check_goal
//                     (:492-512), g/h (:833-847), duplicate test against the
//                     visited grid (:850-854)
//   k_grid_*          contender grid of FREE, non-duplicate final states
//   k_resolve         ordered (ticketed) sequential-equivalent duplicate and
//                     branching-cap resolution (:856-918)
#include <hip/hip_runtime.h>
#include <limits.h>

#include "dp_internal.h"

namespace {

constexpr int kWaves = 4;  // waves per workgroup of the check kernels

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int lane) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

// ------------------------------------------------------------------------
// small utility kernels (no hipMemset on the expansion stream)
// ------------------------------------------------------------------------
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) p[t] = v;
}

__global__ void k_zero_counters(long long* c, int n) {
  if (threadIdx.x < n) c[threadIdx.x] = 0;
}

// In-place exclusive scan of data[0..n-1] into data[0..n] (data[n] = total).
// One workgroup of 1024; each thread owns a contiguous chunk.
__global__ __launch_bounds__(1024) void k_scan(int32_t* data, int64_t n, int32_t* total, int64_t cap,
                                               long long* counters, int count_slot) {
  __shared__ long long s[1024];
  const int t = threadIdx.x;
  const int64_t chunk = (n + 1023) / 1024;
  const int64_t b = t * chunk;
  const int64_t e = b + chunk < n ? b + chunk : n;
  long long sum = 0;
  for (int64_t k = b; k < e; ++k) sum += data[k];
  s[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    long long v = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  long long run = s[t] - sum;
  for (int64_t k = b; k < e; ++k) {
    int32_t v = data[k];
    data[k] = (int32_t)run;
    run += v;
  }
  if (t == 1023) {
    long long tot = s[1023];
    data[n] = (int32_t)tot;
    if (total) *total = (int32_t)(tot < cap ? tot : cap);
    if (count_slot >= 0) counters[count_slot] = tot;
    if (tot > cap) counters[DPC_OVERFLOW] = 1;
  }
}

// ------------------------------------------------------------------------
// applicability (T_m->nearestR over primitive start states)
// ------------------------------------------------------------------------
template <int NX, bool FILL>
__global__ __launch_bounds__(64) void k_app(const dp_dev_sys* __restrict__ Sp, const double* __restrict__ nodes,
                                            int32_t Ncap, int32_t N, const double* __restrict__ start, int32_t P,
                                            double r_app, int32_t* __restrict__ app, int32_t* __restrict__ pnode,
                                            int32_t* __restrict__ pprim, int64_t pair_cap) {
  const dp_dev_sys& S = *Sp;
  const int32_t i = blockIdx.x;
  if (i >= N) return;
  const int lane = threadIdx.x;
  double a[NX];
#pragma unroll
  for (int c = 0; c < NX; ++c) a[c] = dp_is_offset_dim(S, c) ? 0.0 : nodes[(int64_t)c * Ncap + i];
  int64_t base = FILL ? (int64_t)app[i] : 0;
  int32_t cnt = 0;
  for (int32_t j0 = 0; j0 < P; j0 += 64) {
    const int32_t j = j0 + lane;
    bool hit = false;
    if (j < P) {
      double b[NX];
#pragma unroll
      for (int c = 0; c < NX; ++c) b[c] = start[(int64_t)c * P + j];
      hit = dp_distance(S, a, b) <= r_app;
    }
    const uint64_t m = __ballot(hit);
    if (FILL && hit) {
      const int64_t q = base + cnt + __popcll(m & ((1ull << lane) - 1ull));
      if (q < pair_cap) {
        pnode[q] = i;
        pprim[q] = j;
      }
    }
    cnt += __popcll(m);
  }
  if (!FILL && lane == 0) app[i] = cnt;
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
  int32_t P;
  const int32_t* goalidx;
  int32_t num_goal, check_goal;
  const double* obs;
  int32_t nobs;
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
  int32_t* res;
  int8_t* flags;
  double* fin;
  int64_t pcap;
  double* gout;
  double* hout;
  double goal[DP_MAX_NX];
  double goal_bound, r_dup, cdf;
  long long* counters;
};

namespace {

template <int NX>
__device__ __forceinline__ void load_translated(const dp_check_args& a, int64_t sidx, double ox, double oy, double* s) {
  s[0] = a.st[sidx] + ox;
  s[1] = a.st[(int64_t)a.S + sidx] + oy;
#pragma unroll
  for (int c = 2; c < NX; ++c) s[c] = a.st[(int64_t)c * a.S + sidx];
}

// Any visited state within r of x?  The calling lane scans cell `cq` of the
// 3^pos_dim neighbourhood of x's cell.
template <int NX>
__device__ __forceinline__ bool grid_cell_any(const dp_dev_sys& S, const dp_grid_geom& G, const int32_t* start,
                                              const double* pts, int64_t stride, const double* x, int cq, double r) {
  int d0 = cq % 3 - 1, d1 = (cq / 3) % 3 - 1, d2 = cq / 9 - 1;
  if (G.pos_dim < 3) d2 = 0;
  int32_t c0 = dp_cell_axis(x[0], G.lo[0], G.inv_cell, G.dims[0]) + d0;
  int32_t c1 = dp_cell_axis(x[1], G.lo[1], G.inv_cell, G.dims[1]) + d1;
  int32_t c2 = G.pos_dim > 2 ? dp_cell_axis(x[2], G.lo[2], G.inv_cell, G.dims[2]) + d2 : 0;
  if (c0 < 0 || c0 >= G.dims[0] || c1 < 0 || c1 >= G.dims[1] || c2 < 0 || c2 >= G.dims[2]) return false;
  const int64_t cell = ((int64_t)c2 * G.dims[1] + c1) * G.dims[0] + c0;
  const int32_t e = start[cell + 1];
  for (int32_t p = start[cell]; p < e; ++p) {
    double y[NX];
#pragma unroll
    for (int c = 0; c < NX; ++c) y[c] = pts[(int64_t)c * stride + p];
    if (dp_distance(S, x, y) <= r) return true;
  }
  return false;
}

// Planar translate systems (unicycle1/2, car+trailer), FP64.
template <int NX, int LPP>
__global__ __launch_bounds__(256) void k_check_planar(dp_check_args a) {
  extern __shared__ double smem[];
  const dp_dev_sys& S = *a.sys;
  const int nobs = a.nobs;
  if ((int64_t)blockIdx.x * kWaves * (64 / LPP) >= (int64_t)*a.total) return;  // uniform per block
  double* s_obs = smem;  // [nobs][4]
  int16_t* s_cand = (int16_t*)(smem + 4 * nobs);
  for (int t = threadIdx.x; t < 4 * nobs; t += blockDim.x) s_obs[t] = a.obs[t];
  __syncthreads();

  constexpr int SEG = 64 / LPP;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int seg = lane / LPP, k0 = lane % LPP;
  const int shift = seg * LPP;
  const uint64_t segbits = LPP == 64 ? ~0ull : ((1ull << LPP) - 1ull);
  const uint64_t ltmask = (1ull << k0) - 1ull;
  int16_t* cand = s_cand + (int64_t)(wave * SEG + seg) * nobs;
  const int32_t total = *a.total;
  const int ncq = a.grid.pos_dim == 3 ? 27 : 9;
  long long c_b = 0, c_c = 0, c_f = 0, c_g = 0, c_k = 0;

  for (int64_t q0 = ((int64_t)blockIdx.x * kWaves + wave) * SEG; q0 < total;
       q0 += (int64_t)gridDim.x * kWaves * SEG) {
    const int64_t q = q0 + seg;
    const bool vp = q < total;
    int32_t i = 0, j = 0, st0 = 0, L1 = 0;
    double ox = 0.0, oy = 0.0;
    if (vp) {
      i = a.pnode[q];
      j = a.pprim[q];
      st0 = a.off[j];
      L1 = a.off[j + 1] - st0;
      ox = a.nodes[i];
      oy = a.nodes[(int64_t)a.Ncap + i];
    }
    // ---- is_state_valid on every translated state -----------------------
    int first_bad = INT_MAX;
    for (int kb = 0; __ballot(vp && kb < L1) != 0; kb += LPP) {
      const int k = kb + k0;
      bool bad = false;
      if (vp && k < L1) {
        const int64_t sidx = (int64_t)st0 + k;
        double s[NX];
        load_translated<NX>(a, sidx, ox, oy, s);
#pragma unroll
        for (int c = 0; c < NX; ++c) bad = bad || !(s[c] >= S.lo[c] && s[c] <= S.hi[c]);
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
      int ncand = 0;
      for (int ob = 0; ob < nobs; ob += LPP) {
        const int o = ob + k0;
        bool hit = false;
        if (need_col && o < nobs) {
          const double* b = s_obs + 4 * o;
          hit = !(b[0] + b[2] < bx0 || b[0] - b[2] > bx1 || b[1] + b[3] < by0 || b[1] - b[3] > by1);
        }
        const uint64_t m = (__ballot(hit) >> shift) & segbits;
        if (hit) cand[ncand + __popcll(m & ltmask)] = (int16_t)o;
        ncand += __popcll(m);
      }
      wave_lds_sync();
      for (int kb = 0; __ballot(need_col && kb < L1) != 0; kb += LPP) {
        const int k = kb + k0;
        bool col = false;
        if (need_col && k < L1 && first_col == INT_MAX) {
          const int64_t sidx = (int64_t)st0 + k;
          const double x = a.st[sidx] + ox;
          const double y = a.st[(int64_t)a.S + sidx] + oy;
          for (int bi = 0; bi < a.nbox && !col; ++bi) {
            const double c = a.geo[(int64_t)(4 * bi + 2) * a.S + sidx];
            const double s = a.geo[(int64_t)(4 * bi + 3) * a.S + sidx];
            double cx = x, cy = y, hl = S.body_hl, hw = S.body_hw;
            if (bi > 0) {
              cx = x + a.geo[(int64_t)(4 * bi) * a.S + sidx];
              cy = y + a.geo[(int64_t)(4 * bi + 1) * a.S + sidx];
              hl = S.trailer_hl;
              hw = S.trailer_hw;
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

    // ---- FREE: check_goal, g, h, duplicate vs visited --------------------
    const bool fr = vp && vd == DP_V_FREE;
    int chosen = -1;
    if (a.check_goal && __ballot(fr)) {
      bool hit = false;
      if (fr && k0 < a.num_goal) {
        const int idx = a.goalidx[(int64_t)k0 * a.P + j];
        double s[NX];
        load_translated<NX>(a, (int64_t)st0 + idx, ox, oy, s);
        hit = dp_distance(S, s, a.goal) <= a.goal_bound;
      }
      const uint64_t m = (__ballot(hit) >> shift) & segbits;
      if (fr && m) chosen = a.goalidx[(int64_t)(__ffsll((long long)m) - 1) * a.P + j];
    }
    double xf[NX];
    double gq = 0.0, hq = 0.0;
    if (fr) {
      const int fi = chosen != -1 ? chosen : L1 - 1;
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
    if (a.nv > 0 && __ballot(fr)) {
      bool hit = false;
      for (int cb = 0; cb < ncq; cb += LPP) {
        const int cq = cb + k0;
        if (fr && cq < ncq && !hit) hit = grid_cell_any<NX>(S, a.grid, a.vstart, a.vis, a.vstride, xf, cq, a.r_dup);
      }
      dupv = ((__ballot(hit) >> shift) & segbits) != 0;
    }
    if (vp && k0 == 0) {
      a.verdict[q] = vd;
      a.firstk[q] = vd == DP_V_BOUNDS ? first_bad : (vd == DP_V_COLLIDE ? first_col : -1);
      a.goal_out[q] = chosen;
      a.res[q] = fr ? DP_R_UNDECIDED : DP_R_NONE;
      a.flags[q] = (int8_t)((dupv ? 1 : 0) | (chosen != -1 ? 2 : 0));
      if (fr) {
#pragma unroll
        for (int c = 0; c < NX; ++c) a.fin[(int64_t)c * a.pcap + q] = xf[c];
        a.gout[q] = gq;
        a.hout[q] = hq;
      }
      c_b += vd == DP_V_BOUNDS;
      c_c += vd == DP_V_COLLIDE;
      c_f += fr;
      c_g += fr && chosen != -1;
      c_k += vd != DP_V_BOUNDS;
    }
  }
  if (k0 == 0) {
    if (c_b) atomicAdd((unsigned long long*)&a.counters[DPC_BOUNDS], (unsigned long long)c_b);
    if (c_c) atomicAdd((unsigned long long*)&a.counters[DPC_COLLIDE], (unsigned long long)c_c);
    if (c_f) atomicAdd((unsigned long long*)&a.counters[DPC_FREE], (unsigned long long)c_f);
    if (c_g) atomicAdd((unsigned long long*)&a.counters[DPC_GOAL], (unsigned long long)c_g);
    if (c_k) atomicAdd((unsigned long long*)&a.counters[DPC_CHECKED], (unsigned long long)c_k);
  }
}

// ------------------------------------------------------------------------
// grids (visited set and batch contenders): counting sort by cell
// ------------------------------------------------------------------------
template <int NX>
__device__ __forceinline__ int64_t cell_of(const dp_grid_geom& G, const double* pts, int64_t stride, int64_t p) {
  int32_t c0 = dp_cell_axis(pts[p], G.lo[0], G.inv_cell, G.dims[0]);
  int32_t c1 = dp_cell_axis(pts[stride + p], G.lo[1], G.inv_cell, G.dims[1]);
  int32_t c2 = G.pos_dim > 2 ? dp_cell_axis(pts[2 * stride + p], G.lo[2], G.inv_cell, G.dims[2]) : 0;
  return ((int64_t)c2 * G.dims[1] + c1) * G.dims[0] + c0;
}

template <int NX>
__global__ void k_vis_count(dp_grid_geom G, const double* __restrict__ pts, int64_t stride, int64_t n,
                            int32_t* __restrict__ cellof, int32_t* __restrict__ counts) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    int64_t c = cell_of<NX>(G, pts, stride, p);
    cellof[p] = (int32_t)c;
    atomicAdd(&counts[c], 1);
  }
}

template <int NX>
__global__ void k_vis_scatter(const double* __restrict__ pts, int64_t stride, int64_t n,
                              const int32_t* __restrict__ cellof, const int32_t* __restrict__ start,
                              int32_t* __restrict__ cursor, double* __restrict__ out) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = cellof[p];
    const int32_t pos = start[c] + atomicAdd(&cursor[c], 1);
#pragma unroll
    for (int k = 0; k < NX; ++k) out[(int64_t)k * stride + pos] = pts[(int64_t)k * stride + p];
  }
}

// Contenders: FREE pairs that may become NOVEL (goal-forced, or not a
// duplicate of the visited set).
__device__ __forceinline__ bool is_contender(const int8_t* verdict, const int8_t* flags, int64_t q) {
  if (verdict[q] != DP_V_FREE) return false;
  const int8_t f = flags[q];
  return (f & 2) || !(f & 1);
}

template <int NX>
__global__ void k_cont_count(dp_grid_geom G, const int32_t* __restrict__ total, const int8_t* __restrict__ verdict,
                             const int8_t* __restrict__ flags, const double* __restrict__ fin, int64_t pcap,
                             int32_t* __restrict__ counts) {
  const int64_t n = *total;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    if (!is_contender(verdict, flags, q)) continue;
    atomicAdd(&counts[cell_of<NX>(G, fin, pcap, q)], 1);
  }
}

template <int NX>
__global__ void k_cont_scatter(dp_grid_geom G, const int32_t* __restrict__ total, const int8_t* __restrict__ verdict,
                               const int8_t* __restrict__ flags, const double* __restrict__ fin, int64_t pcap,
                               const int32_t* __restrict__ start, int32_t* __restrict__ cursor,
                               int32_t* __restrict__ idx) {
  const int64_t n = *total;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    if (!is_contender(verdict, flags, q)) continue;
    const int64_t c = cell_of<NX>(G, fin, pcap, q);
    idx[start[c] + atomicAdd(&cursor[c], 1)] = (int32_t)q;
  }
}

}  // namespace

struct dp_resolve_args {
  const dp_dev_sys* sys;
  int32_t N;
  const int32_t* app;  // pair offsets per node [N+1]
  const int32_t* total;
  const int8_t* verdict;
  const int8_t* flags;
  const double* fin;
  int64_t pcap;
  int32_t* res;
  dp_grid_geom grid;
  const int32_t* cstart;
  const int32_t* cidx;
  double r_dup;
  int32_t limit;
  int32_t* node_novel;  // [N] number of NOVEL pairs per node (may be null)
  long long* counters;
};

namespace {

// Ordered resolution.  Waves take nodes in ticket order; a wave only ever
// waits for pairs of nodes with smaller tickets, which are held by waves that
// are already resident, so the smallest unfinished ticket never waits.
template <int NX>
__global__ __launch_bounds__(64) void k_resolve(dp_resolve_args a) {
  __shared__ int32_t s_node;
  const dp_dev_sys& S = *a.sys;
  const int lane = threadIdx.x;
  if (lane == 0) s_node = (int32_t)atomicAdd((unsigned long long*)&a.counters[DPC_TICKET], 1ull);
  __syncthreads();
  const int32_t node = s_node;
  if (node >= a.N) return;
  const int32_t tot = *a.total;
  const int32_t begin = a.app[node] < tot ? a.app[node] : tot;
  const int32_t end = a.app[node + 1] < tot ? a.app[node + 1] : tot;
  const int ncq = a.grid.pos_dim == 3 ? 27 : 9;
  int32_t count = 0;
  long long spins = 0, n_nov = 0, n_cap = 0, n_dv = 0, n_db = 0;
  bool err = false;

  for (int32_t cb = begin; cb < end; cb += 64) {
    const int32_t p = cb + lane;
    const bool act = p < end;
    const bool fr = act && a.verdict[p] == DP_V_FREE;
    const int8_t fl = fr ? a.flags[p] : 0;
    const bool forced = (fl & 2) != 0;
    const bool dupv = (fl & 1) != 0;
    const bool checker = fr && !forced && !dupv;
    bool ext = false;
    uint64_t same = 0;
    if (checker) {
      double x[NX];
#pragma unroll
      for (int c = 0; c < NX; ++c) x[c] = a.fin[(int64_t)c * a.pcap + p];
      int32_t cc[3];
      cc[0] = dp_cell_axis(x[0], a.grid.lo[0], a.grid.inv_cell, a.grid.dims[0]);
      cc[1] = dp_cell_axis(x[1], a.grid.lo[1], a.grid.inv_cell, a.grid.dims[1]);
      cc[2] = a.grid.pos_dim > 2 ? dp_cell_axis(x[2], a.grid.lo[2], a.grid.inv_cell, a.grid.dims[2]) : 0;
      for (int cq = 0; cq < ncq && !ext; ++cq) {
        const int32_t c0 = cc[0] + cq % 3 - 1, c1 = cc[1] + (cq / 3) % 3 - 1;
        const int32_t c2 = a.grid.pos_dim > 2 ? cc[2] + cq / 9 - 1 : 0;
        if (c0 < 0 || c0 >= a.grid.dims[0] || c1 < 0 || c1 >= a.grid.dims[1] || c2 < 0 || c2 >= a.grid.dims[2])
          continue;
        const int64_t cell = ((int64_t)c2 * a.grid.dims[1] + c1) * a.grid.dims[0] + c0;
        const int32_t e = a.cstart[cell + 1];
        for (int32_t t = a.cstart[cell]; t < e && !ext; ++t) {
          const int32_t o = a.cidx[t];
          if (o >= p) continue;
          double y[NX];
#pragma unroll
          for (int c = 0; c < NX; ++c) y[c] = a.fin[(int64_t)c * a.pcap + o];
          if (!(dp_distance(S, x, y) <= a.r_dup)) continue;
          if (o >= cb) {
            same |= 1ull << (o - cb);
            continue;
          }
          int r = __hip_atomic_load(&a.res[o], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          int guard = 0;
          while (r == DP_R_UNDECIDED && guard < DP_SPIN_LIMIT) {
            __builtin_amdgcn_s_sleep(1);
            ++guard;
            r = __hip_atomic_load(&a.res[o], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          }
          spins += guard;
          if (r == DP_R_UNDECIDED) err = true;
          if (r == DP_R_NOVEL) ext = true;
        }
      }
    }
    const uint64_t free_m = __ballot(fr);
    const uint64_t forced_m = __ballot(fr && forced);
    const uint64_t dupv_m = __ballot(fr && dupv && !forced);
    const uint64_t ext_m = __ballot(checker && ext);
    uint64_t novel_m = 0, capped_m = 0, dupb_m = 0, dupvr_m = 0;
    uint64_t rem = free_m;
    while (rem) {
      const int b = __ffsll((long long)rem) - 1;
      rem &= rem - 1ull;
      const uint64_t bit = 1ull << b;
      if (a.limit > 0 && count + __popcll(novel_m) >= a.limit) {
        capped_m |= bit;
      } else if (forced_m & bit) {
        novel_m |= bit;
      } else if (dupv_m & bit) {
        dupvr_m |= bit;
      } else if (ext_m & bit) {
        dupb_m |= bit;
      } else if (readlane64(same, b) & novel_m) {
        dupb_m |= bit;
      } else {
        novel_m |= bit;
      }
    }
    if (fr) {
      const uint64_t bit = 1ull << lane;
      const int r = (capped_m & bit) ? DP_R_CAPPED
                    : (novel_m & bit) ? DP_R_NOVEL
                    : (dupvr_m & bit) ? DP_R_DUP_VISITED
                                      : DP_R_DUP_BATCH;
      __hip_atomic_store(&a.res[p], r, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    count += __popcll(novel_m);
    n_nov += __popcll(novel_m);
    n_cap += __popcll(capped_m);
    n_dv += __popcll(dupvr_m);
    n_db += __popcll(dupb_m);
  }
  // reduce spins / err across the wave
  for (int o = 32; o > 0; o >>= 1) spins += __shfl_xor(spins, o);
  const bool any_err = __ballot(err) != 0;
  if (lane == 0) {
    if (a.node_novel) a.node_novel[node] = count;
    if (n_nov) atomicAdd((unsigned long long*)&a.counters[DPC_NOVEL], (unsigned long long)n_nov);
    if (n_cap) atomicAdd((unsigned long long*)&a.counters[DPC_CAPPED], (unsigned long long)n_cap);
    if (n_dv) atomicAdd((unsigned long long*)&a.counters[DPC_DUPV], (unsigned long long)n_dv);
    if (n_db) atomicAdd((unsigned long long*)&a.counters[DPC_DUPB], (unsigned long long)n_db);
    if (spins) atomicAdd((unsigned long long*)&a.counters[DPC_SPINS], (unsigned long long)spins);
    if (any_err) atomicOr((unsigned long long*)&a.counters[DPC_ERRFLAG], 1ull);
  }
}

// Append the NOVEL final states of the batch to the visited set in canonical
// order: node_novel was exclusive-scanned into per-node destinations.
template <int NX>
__global__ __launch_bounds__(64) void k_gather_novel(int32_t N, const int32_t* __restrict__ app,
                                                     const int32_t* __restrict__ total,
                                                     const int32_t* __restrict__ res, const int32_t* __restrict__ dst,
                                                     const double* __restrict__ fin, int64_t pcap,
                                                     double* __restrict__ vis, int64_t vstride, int64_t vbase) {
  const int32_t node = blockIdx.x;
  if (node >= N) return;
  const int lane = threadIdx.x;
  const int32_t tot = *total;
  const int32_t begin = app[node] < tot ? app[node] : tot;
  const int32_t end = app[node + 1] < tot ? app[node + 1] : tot;
  int64_t w = vbase + dst[node];
  for (int32_t cb = begin; cb < end; cb += 64) {
    const int32_t p = cb + lane;
    const bool nov = p < end && res[p] == DP_R_NOVEL;
    const uint64_t m = __ballot(nov);
    if (nov) {
      const int64_t pos = w + __popcll(m & ((1ull << lane) - 1ull));
#pragma unroll
      for (int c = 0; c < NX; ++c) vis[(int64_t)c * vstride + pos] = fin[(int64_t)c * pcap + p];
    }
    w += __popcll(m);
  }
}

}  // namespace

// ==========================================================================
// host-side launchers
// ==========================================================================

#define DP_HIP_RET(expr)                                                        \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      dp_set_error(std::string(#expr) + ": " + hipGetErrorString(e_));          \
      return DP_ERR_HIP;                                                        \
    }                                                                           \
  } while (0)

static int grid_blocks(int64_t work, int per_block, int max_blocks) {
  int64_t b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (int)b;
}

// Grid geometry for radius r over the world bounds (cells of r/w0, padded).
static dp_grid_geom make_grid(const dp_ctx* c, double r) {
  dp_grid_geom G{};
  G.pos_dim = c->desc.pos_dim;
  double cell = (r / c->desc.dist_w[0]) * (1.0 + 1e-9);
  const int64_t max_cells = 1ll << 24;
  for (;;) {
    int64_t n = 1;
    for (int d = 0; d < 3; ++d) {
      if (d < G.pos_dim) {
        double ext = c->world_hi[d] - c->world_lo[d];
        int64_t k = (int64_t)(ext / cell) + 3;
        G.dims[d] = (int32_t)k;
        G.lo[d] = c->world_lo[d] - cell;
      } else {
        G.dims[d] = 1;
        G.lo[d] = 0.0;
      }
      n *= G.dims[d];
    }
    if (n <= max_cells) {
      G.ncells = n;
      break;
    }
    cell *= 1.5;  // coarser cells stay exact, only slower
  }
  G.inv_cell = 1.0 / cell;
  return G;
}

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

static int ensure_i32(int32_t** p, int64_t* cap, int64_t need) {
  if (*cap >= need) return DP_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  DP_HIP_RET(hipMalloc(p, sizeof(int32_t) * need));
  *cap = need;
  return DP_OK;
}

int dp_launch_visited_rebuild(dp_ctx* c, double radius, hipStream_t s) {
  c->grid = make_grid(c, radius);
  c->grid_radius = radius;
  const int64_t nc = c->grid.ncells;
  int rc = ensure_i32(&c->d_vis_start, &c->vis_cells_cap, nc + 1);
  if (rc) return rc;
  if (c->nv == 0) {
    k_fill_i32<<<grid_blocks(nc + 1, 256, 4096), 256, 0, s>>>(c->d_vis_start, nc + 1, 0);
    c->grid_dirty = false;
    return DP_OK;
  }
  k_fill_i32<<<grid_blocks(nc + 1, 256, 4096), 256, 0, s>>>(c->d_vis_start, nc + 1, 0);
  const int nx = c->desc.nx;
  rc = dispatch_nx(nx, [&](auto NXc) -> int {
    constexpr int NX = decltype(NXc)::value;
    k_vis_count<NX><<<grid_blocks(c->nv, 256, 8192), 256, 0, s>>>(c->grid, c->d_vis, c->nv_cap, c->nv,
                                                                   c->d_vis_cell, c->d_vis_start);
    return DP_OK;
  });
  if (rc) return rc;
  k_scan<<<1, 1024, 0, s>>>(c->d_vis_start, nc, nullptr, INT64_MAX, c->d_counters, -1);
  // cursor: reuse the contender cursor buffer
  rc = ensure_i32(&c->d_cont_cursor, &c->cont_cursor_cap, nc + 1);
  if (rc) return rc;
  k_fill_i32<<<grid_blocks(nc, 256, 4096), 256, 0, s>>>(c->d_cont_cursor, nc, 0);
  rc = dispatch_nx(nx, [&](auto NXc) -> int {
    constexpr int NX = decltype(NXc)::value;
    k_vis_scatter<NX><<<grid_blocks(c->nv, 256, 8192), 256, 0, s>>>(c->d_vis, c->nv_cap, c->nv, c->d_vis_cell,
                                                                     c->d_vis_start, c->d_cont_cursor,
                                                                     c->d_vis_sorted);
    return DP_OK;
  });
  if (rc) return rc;
  DP_HIP_RET(hipGetLastError());
  c->grid_dirty = false;
  return DP_OK;
}

int dp_launch_expand(dp_ctx* c, const dp_expand_params* p, hipStream_t s) {
  const int nx = c->desc.nx;
  const int32_t N = c->N;
  const double r_app = (double)(p->alpha * p->delta);
  const double r_dup = (1. - (double)p->alpha) * (double)p->delta;
  const double goal_bound = p->delta_factor_goal * (double)p->delta;

  if (c->grid_dirty || c->grid_radius != r_dup) {
    int rc = dp_launch_visited_rebuild(c, r_dup, s);
    if (rc) return rc;
  }
  const int64_t nc = c->grid.ncells;

  DP_HIP_RET(hipEventRecord(c->ev[0], s));
  k_zero_counters<<<1, 64, 0, s>>>(c->d_counters, DPC_COUNT);
  int rc = dispatch_nx(nx, [&](auto NXc) -> int {
    constexpr int NX = decltype(NXc)::value;
    k_app<NX, false><<<N, 64, 0, s>>>(c->d_sys, c->d_nodes, c->N_cap, N, c->d_start, c->P, r_app, c->d_app,
                                     nullptr, nullptr, 0);
    k_scan<<<1, 1024, 0, s>>>(c->d_app, N, c->d_total, c->pair_cap, c->d_counters, DPC_PAIRS);
    k_app<NX, true><<<N, 64, 0, s>>>(c->d_sys, c->d_nodes, c->N_cap, N, c->d_start, c->P, r_app, c->d_app,
                                    c->d_pnode, c->d_pprim, c->pair_cap);
    return DP_OK;
  });
  if (rc) return rc;
  DP_HIP_RET(hipEventRecord(c->ev[1], s));

  // ---- per-pair verdicts ---------------------------------------------------
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
  a.P = c->P;
  a.goalidx = c->d_goalidx;
  a.num_goal = p->num_check_goal;
  a.check_goal = p->check_intermediate_goal && p->num_check_goal > 0;
  a.obs = c->d_obs;
  a.nobs = c->nb;
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
  a.counters = c->d_counters;

  if (c->desc.system == DP_SYS_QUADROTOR) {
    dp_set_error("quadrotor expansion kernel not available in this build");
    return DP_ERR_ARG;
  }
  const int lpp = c->maxlen <= 16 ? 16 : (c->maxlen <= 32 ? 32 : 64);
  const int seg = 64 / lpp;
  const size_t lds = sizeof(double) * 4 * c->nb + sizeof(int16_t) * (size_t)kWaves * seg * (c->nb > 0 ? c->nb : 1);
  // The pair count is on the device; size the grid from the bound N*P.
  const int64_t groups = ((int64_t)N * c->P + seg - 1) / seg;
  const int blocks = grid_blocks(groups, kWaves, 2048);
  rc = dispatch_nx(nx, [&](auto NXc) -> int {
    constexpr int NX = decltype(NXc)::value;
    if constexpr (NX == 13) {
      return DP_ERR_ARG;
    } else {
      if (lpp == 16) k_check_planar<NX, 16><<<blocks, 256, lds, s>>>(a);
      else if (lpp == 32) k_check_planar<NX, 32><<<blocks, 256, lds, s>>>(a);
      else k_check_planar<NX, 64><<<blocks, 256, lds, s>>>(a);
      return DP_OK;
    }
  });
  if (rc) return rc;
  DP_HIP_RET(hipGetLastError());
  DP_HIP_RET(hipEventRecord(c->ev[2], s));

  // ---- batch duplicate resolution -----------------------------------------
  rc = ensure_i32(&c->d_cont_start, &c->cont_start_cap, nc + 1);
  if (rc) return rc;
  rc = ensure_i32(&c->d_cont_cursor, &c->cont_cursor_cap, nc + 1);
  if (rc) return rc;
  rc = dispatch_nx(nx, [&](auto NXc) -> int {
    constexpr int NX = decltype(NXc)::value;
    const int pb = grid_blocks(c->pair_cap, 256, 8192);
    k_fill_i32<<<grid_blocks(nc + 1, 256, 4096), 256, 0, s>>>(c->d_cont_start, nc + 1, 0);
    k_cont_count<NX><<<pb, 256, 0, s>>>(c->grid, c->d_total, c->d_verdict, c->d_flags, c->d_final, c->pair_cap,
                                        c->d_cont_start);
    k_scan<<<1, 1024, 0, s>>>(c->d_cont_start, nc, nullptr, INT64_MAX, c->d_counters, -1);
    k_fill_i32<<<grid_blocks(nc, 256, 4096), 256, 0, s>>>(c->d_cont_cursor, nc, 0);
    k_cont_scatter<NX><<<pb, 256, 0, s>>>(c->grid, c->d_total, c->d_verdict, c->d_flags, c->d_final, c->pair_cap,
                                          c->d_cont_start, c->d_cont_cursor, c->d_cont_idx);
    dp_resolve_args r{};
    r.sys = c->d_sys;
    r.N = N;
    r.app = c->d_app;
    r.total = c->d_total;
    r.verdict = c->d_verdict;
    r.flags = c->d_flags;
    r.fin = c->d_final;
    r.pcap = c->pair_cap;
    r.res = c->d_res;
    r.grid = c->grid;
    r.cstart = c->d_cont_start;
    r.cidx = c->d_cont_idx;
    r.r_dup = r_dup;
    r.limit = p->limit_branching_factor;
    r.node_novel = c->d_node_novel;
    r.counters = c->d_counters;
    k_resolve<NX><<<N, 64, 0, s>>>(r);
    return DP_OK;
  });
  if (rc) return rc;
  DP_HIP_RET(hipGetLastError());
  DP_HIP_RET(hipEventRecord(c->ev[3], s));
  c->timing_valid = true;
  return DP_OK;
}

int dp_launch_visited_append_novel(dp_ctx* c, int64_t novel_total, hipStream_t s) {
  // Caller guarantees visited capacity >= nv + novel_total.
  const int32_t N = c->N;
  if (N <= 0 || novel_total <= 0) return DP_OK;
  k_scan<<<1, 1024, 0, s>>>(c->d_node_novel, N, nullptr, INT64_MAX, c->d_counters, -1);
  int rc = dispatch_nx(c->desc.nx, [&](auto NXc) -> int {
    constexpr int NX = decltype(NXc)::value;
    k_gather_novel<NX><<<N, 64, 0, s>>>(N, c->d_app, c->d_total, c->d_res, c->d_node_novel, c->d_final,
                                        c->pair_cap, c->d_vis, c->nv_cap, c->nv);
    return DP_OK;
  });
  if (rc) return rc;
  DP_HIP_RET(hipGetLastError());
  c->nv += novel_total;
  c->grid_dirty = true;
  return DP_OK;
}
