// k_grid.hip -- duplicate tests of the expansion: the visited-set grid (T_n),
// the in-batch contender grid, earlier-conflict lists and the resolution that
// reproduces the reference's sequential insert semantics
// (dbastar.cpp:850-918) for a whole batch.
//
// Sequential semantics restated: in canonical (node, primitive) order, a FREE
// pair of node i is CAPPED once limit NOVEL pairs of node i precede it;
// otherwise it is NOVEL when goal-forced (chosen_index != -1) or when no
// visited state and no earlier NOVEL pair lies within (1-alpha)*delta;
// otherwise it is a duplicate.
//
// Parallel form: each FREE pair that could become NOVEL ("contender") gets
// the list of earlier contenders within the radius (k_conf_waves).  Statuses
// that are certain from the lists alone are published first (k_pre_nodes).
// Then one wave per node, claimed in node order by a ticket, walks its pairs
// 64 at a time (k_resolve_ticketed); a pair is decided as soon as every
// earlier conflicting contender is decided (or one is NOVEL).  The smallest
// undecided pair of the batch is always decidable, so the walk terminates;
// every wait is bounded, and a host-driven round kernel (k_round) finishes
// what a wave gives up on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dp_internal.h"

// Cell sizes: at least r/w0 on the position axes and r/w1 on the heading
// axis, never narrower than the largest world extent / 2^20 (a radius of 0,
// or a tiny one, would otherwise give empty cells; coarser cells stay
// exact).  Counts are formed in double and coarsened until the grid holds at
// most 2^24 cells.
dp_grid_geom dp_make_grid(const dp_ctx* c, double r) {
  dp_grid_geom G{};
  const bool planar = c->desc.pos_dim == 2;
  const int npos = planar ? 2 : 3;
  double ext_max = 0.0;
  for (int d = 0; d < npos; ++d) ext_max = std::max(ext_max, c->world_hi[d] - c->world_lo[d]);
  const double min_cell = ext_max > 0.0 ? ext_max / (double)(1 << 20) : 1.0;
  double cell = (r / c->desc.dist_w[0]) * (1.0 + 1e-9);
  if (!(cell >= min_cell)) cell = min_cell;
  int32_t nth = 1;
  if (planar && c->desc.dist_w[1] > 0.0) {
    const double bw = (r / c->desc.dist_w[1]) * (1.0 + 1e-6);
    const double k = bw > 0.0 ? std::floor(DP_TWO_PI / bw) : 4096.0;
    nth = k >= 3.0 ? (int32_t)std::min(k, 4096.0) : 1;
  }
  const double max_cells = (double)(1ll << 24);
  for (;;) {
    double n = 1.0;
    for (int d = 0; d < npos; ++d) {
      const double k = std::floor((c->world_hi[d] - c->world_lo[d]) / cell) + 3.0;
      G.dims[d] = (int32_t)std::min(k, (double)INT32_MAX);
      G.lo[d] = c->world_lo[d] - cell;
      G.inv_cell[d] = 1.0 / cell;
      n *= k;
    }
    if (planar) {
      G.dims[2] = nth;
      G.lo[2] = -DP_PI;
      G.inv_cell[2] = (double)nth / DP_TWO_PI;
      G.periodic2 = 1;
      n *= nth;
    }
    if (n <= max_cells) {
      G.ncells = (int64_t)G.dims[0] * G.dims[1] * G.dims[2];
      break;
    }
    cell *= 1.5;  // coarser cells stay exact, only slower
    if (planar && nth > 1) {
      nth = nth / 2;  // wider heading bins stay >= r/w1
      if (nth < 3) nth = 1;
    }
  }
  return G;
}

namespace {

template <int NX>
__device__ __forceinline__ int64_t cell_of_soa(const dp_grid_geom& G, const double* pts, int64_t stride, int64_t p) {
  double x[3] = {pts[p], pts[stride + p], pts[2 * stride + p]};
  int32_t c[3];
  dp_cell_coords(G, x, c);
  return dp_cell_linear(G, c);
}

// ---- visited set ---------------------------------------------------------
template <int NX>
__global__ void k_vis_count(dp_grid_geom G, const double* __restrict__ pts, int64_t stride, int64_t n,
                            int32_t* __restrict__ cellof, int32_t* __restrict__ counts) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = cell_of_soa<NX>(G, pts, stride, p);
    cellof[p] = (int32_t)c;
    atomicAdd(&counts[c], 1);
  }
}

template <int NX>
__global__ void k_vis_scatter(const double* __restrict__ pts, int64_t stride, int64_t n,
                              const int32_t* __restrict__ cellof, const int32_t* __restrict__ start,
                              int32_t* __restrict__ cursor, double* __restrict__ out, int32_t* __restrict__ perm) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = cellof[p];
    const int32_t pos = start[c] + atomicAdd(&cursor[c], 1);
#pragma unroll
    for (int k = 0; k < NX; ++k) out[(int64_t)k * stride + pos] = pts[(int64_t)k * stride + p];
    perm[pos] = (int32_t)p;
  }
}

// ---- contenders --------------------------------------------------------------
__device__ __forceinline__ bool is_contender(int8_t verdict, int8_t flags) {
  return verdict == DP_V_FREE && !(flags & (DP_F_DEFER | DP_F_PRECAP)) && ((flags & 2) || !(flags & 1));
}

template <int NX>
__global__ void k_cont_count(dp_grid_geom G, const int32_t* __restrict__ total, const int8_t* __restrict__ verdict,
                             const int8_t* __restrict__ flags, const double* __restrict__ fin, int64_t pcap,
                             int32_t* __restrict__ counts) {
  const int64_t n = *total;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    if (!is_contender(verdict[q], flags[q])) continue;
    atomicAdd(&counts[cell_of_soa<NX>(G, fin, pcap, q)], 1);
  }
}

template <int NX>
__global__ void k_cont_scatter(dp_grid_geom G, const int32_t* __restrict__ total, const int8_t* __restrict__ verdict,
                               const int8_t* __restrict__ flags, const double* __restrict__ fin, int64_t pcap,
                               const int32_t* __restrict__ start, int32_t* __restrict__ cursor,
                               int32_t* __restrict__ idx, double* __restrict__ xs) {
  const int64_t n = *total;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    if (!is_contender(verdict[q], flags[q])) continue;
    const int64_t c = cell_of_soa<NX>(G, fin, pcap, q);
    const int32_t pos = start[c] + atomicAdd(&cursor[c], 1);
    idx[pos] = (int32_t)q;
#pragma unroll
    for (int k = 0; k < NX; ++k) xs[(int64_t)k * pcap + pos] = fin[(int64_t)k * pcap + q];
  }
}

// Earlier contenders of pair q within r: scans the 27 cells around x.
// Calls f(o) for every contender o < q within r that keep(o) accepts (keep
// runs before the metric and may read memory); stops when f returns false.
// Candidates are taken eight at a time so that their loads overlap.
template <int NX, typename F, typename Keep>
__device__ __forceinline__ void scan_earlier(const dp_dev_sys& S, const dp_grid_geom& G, const int32_t* cstart,
                                             const int32_t* cidx, const double* cx, int64_t pcap, const double* x,
                                             int64_t q, double r, F&& f, Keep&& keep) {
  int32_t c[3];
  dp_cell_coords(G, x, c);
  const double T = dp_pos_bound(S, r);
  for (int cq = 0; cq < 27; ++cq) {
    const int64_t cell = dp_cell_neighbor(G, c, cq);
    if (cell < 0) continue;
    const int32_t e = cstart[cell + 1];
    for (int32_t t0 = cstart[cell]; t0 < e; t0 += 8) {
      int32_t o[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) o[k] = t0 + k < e ? cidx[t0 + k] : INT32_MAX;
      uint32_t km = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) km |= (o[k] < q && keep(o[k]) ? 1u : 0u) << k;
      for (; km; km &= km - 1u) {
        const int32_t t = t0 + __ffs(km) - 1;
        if (!dp_within_col_pf<NX>(S, x, cx, pcap, t, r, T)) continue;
        if (!f(cidx[t])) return;
      }
    }
  }
}

// Earlier-conflict lists, stored inline (up to DP_CONF_SLOTS per pair, pair-major:
// a list's entries are adjacent for its writer and a chunk's lists for its walk)
// for contenders: FREE pairs that are neither goal-forced nor duplicates of
// the visited set (with a window, only those at position < window in their
// node).  Lists longer than DP_CONF_SLOTS continue in the overflow pool;
// conf_cnt = -1 marks a pair without a stored list (past the window, or the
// pool ran out), for which the resolution scans the grid.  The order of a
// list's entries carries no meaning.

// Contenders of one cell that one wave of k_conf_waves takes.  A wave's
// serial work is its contenders times the candidate batches of the block
// (K / batch), so the contenders of a crowded block are split over several
// waves, about 16 contender-batches each, in at most 255 parts.
__device__ __forceinline__ int32_t cell_part_size(int32_t m, int32_t K, int32_t batch) {
  const int32_t nb = (K + batch - 1) / batch;
  int32_t per = nb > 0 ? 16 / nb : 64;
  per = per < 1 ? 1 : (per > 64 ? 64 : per);
  const int32_t least = (m + 254) / 255;
  return per < least ? least : per;
}

// Contenders in the 27-cell block around `cell` (the runs k_conf_waves
// reads, with the same neighbour rules).
__device__ __forceinline__ int32_t cell_block_count(const dp_grid_geom& G, const int32_t* __restrict__ cstart,
                                                    int64_t cell) {
  const int64_t d01 = (int64_t)G.dims[0] * G.dims[1];
  const int32_t c0 = (int32_t)(cell % G.dims[0]), c1 = (int32_t)((cell / G.dims[0]) % G.dims[1]);
  const int32_t c2 = (int32_t)(cell / d01);
  const int32_t x0 = c0 > 0 ? c0 - 1 : 0, x1 = c0 + 1 < G.dims[0] ? c0 + 1 : G.dims[0] - 1;
  int32_t K = 0;
  for (int k = 0; k < 9; ++k) {
    const int32_t n1 = c1 + k % 3 - 1;
    const int32_t d2 = k / 3 - 1;
    int32_t n2 = c2 + d2;
    bool ok = n1 >= 0 && n1 < G.dims[1];
    if (G.dims[2] == 1) {
      ok = ok && d2 == 0;
      n2 = 0;
    } else if (G.periodic2) {
      n2 = n2 < 0 ? n2 + G.dims[2] : (n2 >= G.dims[2] ? n2 - G.dims[2] : n2);
    } else {
      ok = ok && n2 >= 0 && n2 < G.dims[2];
    }
    if (ok) {
      const int64_t row = ((int64_t)n2 * G.dims[1] + n1) * G.dims[0];
      K += cstart[row + x1 + 1] - cstart[row + x0];
    }
  }
  return K;
}

// The non-empty cells of the contender grid, one thread per cell: each
// workgroup reserves room for its listed cells with one global atomic
// (cells are listed in no particular order).  A cell whose 27-cell
// block holds its one contender only is settled here (no earlier
// contender can conflict: an empty list, and NUC unless goal-forced) and is
// not listed, when no window applies.
__global__ __launch_bounds__(256) void k_cell_list(int64_t ncells, dp_grid_geom G, const int32_t* __restrict__ cstart,
                                                   const int32_t* __restrict__ cidx, int32_t window, int32_t batch,
                                                   int8_t* __restrict__ flags, int16_t* __restrict__ conf_cnt,
                                                   int32_t* __restrict__ list, long long* __restrict__ counters) {
  __shared__ int32_t s_cnt;
  __shared__ long long s_base;
  const int lane = threadIdx.x % 64;
  // one cell per thread (the grid covers every cell)
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int kind = 0;  // 0: empty or no list needed, 1: listed, 2: settled here
  int32_t parts = 0;
  if (c < ncells) {
    const int32_t b = cstart[c], e = cstart[c + 1];
    const int32_t K = e > b ? cell_block_count(G, cstart, c) : 0;
    if (e > b) kind = e - b != 1 || K != 1 ? 1 : 2;
    if (kind == 1 && window > 0) {
      // a cell whose contenders are all goal-forced or past their node's
      // list window needs no wave; the latter are scanned when reached
      bool any = false;
      for (int32_t k = b; k < e && !any; ++k) {
        const int8_t fl = flags[cidx[k]];
        any = !(fl & 2) && !(fl & DP_F_NOLIST);
      }
      if (!any) kind = 0;
    }
    if (kind == 1) {
      const int32_t per = cell_part_size(e - b, K, batch);
      parts = (e - b + per - 1) / per;
    }
    if (kind == 2) {
      const int32_t q = cidx[b];
      const int8_t fl = flags[q];
      if (!(fl & 2)) {  // goal-forced pairs are NOVEL without a list
        conf_cnt[q] = 0;
        flags[q] = (int8_t)(fl | DP_F_NUC);
      }
    }
  }
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  // items: cell | part << 24 (a grid has at most 2^24 cells)
  int incl = parts;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int v = __shfl_up(incl, d, 64);
    if (lane >= d) incl += v;
  }
  const int wtot = __shfl(incl, 63, 64);
  int woff = 0;
  if (lane == 0 && wtot) woff = atomicAdd(&s_cnt, wtot);
  __syncthreads();
  if (threadIdx.x == 0)
    s_base = s_cnt ? (long long)atomicAdd((unsigned long long*)&counters[DPC_CELLS], (unsigned long long)s_cnt) : 0;
  __syncthreads();
  woff = __shfl(woff, 0, 64);
  for (int32_t k = 0; k < parts; ++k)
    list[s_base + woff + (incl - parts) + k] = (int32_t)((uint32_t)c | ((uint32_t)k << 24));
}

// One wave per non-empty grid cell: the 27 cells around it
// form 9 runs of 3 cells consecutive in memory (x is the fastest grid
// dimension).  For each of the cell's contenders in turn, the lanes read the
// runs' contenders (coalesced, cached: the same few kilobytes for every
// contender of the cell) and test them, kU per lane in flight, with ballot
// compaction into the list.  No LDS and no barriers, so many cells are in
// flight per CU.
template <int NX>
__global__ __launch_bounds__(256) void k_conf_waves(const dp_dev_sys* __restrict__ Sp, dp_grid_geom G,
                                                    int8_t* __restrict__ flags,
                                                    const int32_t* __restrict__ pnode,
                                                    const int32_t* __restrict__ app, int32_t window, int64_t pcap,
                                                    const int32_t* __restrict__ cstart,
                                                    const int32_t* __restrict__ cidx, const double* __restrict__ cx,
                                                    const int32_t* __restrict__ cells, double r,
                                                    int16_t* __restrict__ conf_cnt, int32_t* __restrict__ conf,
                                                    int32_t* __restrict__ conf_ov, int32_t* __restrict__ pool,
                                                    int64_t pool_cap, int32_t slots, long long* __restrict__ counters) {
  const dp_dev_sys& S = *Sp;
  const int lane = threadIdx.x % 64;
  const uint64_t ltmask = (1ull << lane) - 1ull;
  const long long ncl = counters[DPC_CELLS];
  const double T = dp_pos_bound(S, r);
  // quadrotor: |q.q'| below cos(r / w1) - 1e-9 puts the attitude term alone
  // above r (acos is decreasing with slope <= -1, and the margin dwarfs the
  // series' error), so the pair is rejected without the acos
  const double cq = NX == 13 && S.w[1] > 0.0 && r / S.w[1] < DP_HALF_PI ? cos(r / S.w[1]) - 1e-9 : -1.0;
  const int64_t d01 = (int64_t)G.dims[0] * G.dims[1];
  const long long nwaves = (long long)gridDim.x * (blockDim.x / 64);
  constexpr int kU = NX == 13 ? 2 : 4;  // candidates per lane in flight
  // Workgroups are dealt to the 8 XCDs in turn and each XCD has its own L2;
  // items listed next to each other are cells next to each other, whose
  // 27-cell blocks read mostly the same contender runs, so each XCD takes
  // one eighth of the items, contiguous (when the grid is a multiple of 8
  // workgroups; otherwise every wave strides over all items)
  long long w0 = (long long)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64, w_end = ncl, w_step = nwaves;
  if ((gridDim.x & 7) == 0) {
    const long long range = (ncl + 7) / 8, x = blockIdx.x & 7;
    w0 = x * range + (long long)(blockIdx.x >> 3) * (blockDim.x / 64) + threadIdx.x / 64;
    w_end = x * range + range < ncl ? x * range + range : ncl;
    w_step = nwaves / 8;
  }
  for (long long w = w0; w < w_end; w += w_step) {
    const uint32_t item = (uint32_t)cells[w];
    const int64_t cell = item & 0xFFFFFFu;
    const int32_t part = (int32_t)(item >> 24);
    const int32_t c0 = (int32_t)(cell % G.dims[0]), c1 = (int32_t)((cell / G.dims[0]) % G.dims[1]);
    const int32_t c2 = (int32_t)(cell / d01);
    // run k (lanes 0..8): rows c1 + k%3 - 1, layers c2 + k/3 - 1 (headings, or z for the
    // quadrotor; the rules of dp_cell_neighbor)
    int n = 0, st = 0;
    if (lane < 9) {
      const int32_t n1 = c1 + lane % 3 - 1;
      const int32_t d2 = lane / 3 - 1;
      int32_t n2 = c2 + d2;
      bool ok = n1 >= 0 && n1 < G.dims[1];
      if (G.dims[2] == 1) {
        ok = ok && d2 == 0;
        n2 = 0;
      } else if (G.periodic2) {
        n2 = n2 < 0 ? n2 + G.dims[2] : (n2 >= G.dims[2] ? n2 - G.dims[2] : n2);
      } else {
        ok = ok && n2 >= 0 && n2 < G.dims[2];
      }
      if (ok) {
        const int64_t row = ((int64_t)n2 * G.dims[1] + n1) * G.dims[0];
        const int32_t x0 = c0 > 0 ? c0 - 1 : 0, x1 = c0 + 1 < G.dims[0] ? c0 + 1 : G.dims[0] - 1;
        st = cstart[row + x0];
        n = cstart[row + x1 + 1] - st;
      }
    }
    int incl = n;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
      const int v = __shfl_up(incl, d, 64);
      if (lane >= d) incl += v;
    }
    // run starts and exclusive prefix, wave-uniform
    int32_t rs[9], rp[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      rs[k] = __builtin_amdgcn_readlane(st, k);
      rp[k] = __builtin_amdgcn_readlane(incl - n, k);
    }
    const int K = __builtin_amdgcn_readlane(incl, 8);
    const int32_t b0 = cstart[cell], e0 = cstart[cell + 1];
    const int32_t per = cell_part_size(e0 - b0, K, 64 * kU);
    const int32_t b = b0 + part * per, e = e0 < b + per ? e0 : b + per;
    // the cell's contenders, 64 at a time, one per lane
    for (int32_t g0 = b; g0 < e; g0 += 64) {
      const int m = e - g0 < 64 ? e - g0 : 64;
      int32_t q_l = INT32_MAX;
      bool list_l = false;
      // the group's states, one per lane (broadcast by readlane below)
      double x_l[NX];
#pragma unroll
      for (int k = 0; k < NX; ++k) x_l[k] = 0.0;
      if (lane < m) {
#pragma unroll
        for (int k = 0; k < NX; ++k) x_l[k] = cx[(int64_t)k * pcap + g0 + lane];
        q_l = cidx[g0 + lane];
        const int8_t fl = flags[q_l];
        list_l = !(fl & 2);  // goal-forced: NOVEL without a list
        // past the list window: scanned when reached (k_list_window set
        // its negative list length)
        if (window > 0 && (fl & DP_F_NOLIST)) list_l = false;
      }
      const uint64_t lists = __ballot(list_l);
      int cnt_l = 0;           // list length of this lane's contender so far
      long long off_l = 0;     // its overflow pool offset (pool pass)
      // Candidates in order, kU per lane in flight, each batch tested
      // against every contender of the group.  Entries at list position
      // < DP_CONF_SLOTS go inline; the pool pass writes the others.
      auto pass = [&](bool pool_pass, uint64_t who) {
        for (int j0 = 0; j0 < K; j0 += 64 * kU) {
          // candidate index and state, every component loaded together with
          // the batch (the contender loop below then runs without loads)
          int32_t o[kU];
          int64_t tt[kU];
          double y[kU][NX];
#pragma unroll
          for (int u = 0; u < kU; ++u) {
            const int j = j0 + 64 * u + lane;
            o[u] = INT32_MAX;
            tt[u] = 0;
            if (j < K) {
              int k = 0;
#pragma unroll
              for (int kk = 1; kk < 9; ++kk) k += j >= rp[kk] ? 1 : 0;
              int32_t sk = rs[0], pk = rp[0];
#pragma unroll
              for (int kk = 1; kk < 9; ++kk)
                if (k == kk) sk = rs[kk], pk = rp[kk];
              tt[u] = (int64_t)sk + (j - pk);
              o[u] = cidx[tt[u]];
#pragma unroll
              for (int d = 0; d < NX; ++d) y[u][d] = cx[(int64_t)d * pcap + tt[u]];
            }
          }
          for (uint64_t rem = who; rem; rem &= rem - 1ull) {
            const int c = __ffsll((long long)rem) - 1;
            const int32_t q = __builtin_amdgcn_readlane(q_l, c);
            double x[NX];
#pragma unroll
            for (int k = 0; k < NX; ++k) {
              const uint64_t bits = (uint64_t)__double_as_longlong(x_l[k]);
              const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)bits, c);
              const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(bits >> 32), c);
              x[k] = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
            }
            int cc = __builtin_amdgcn_readlane(cnt_l, c);
            const long long off = pool_pass ? (long long)__builtin_amdgcn_readlane((int)off_l, c) : 0;
#pragma unroll
            for (int u = 0; u < kU; ++u) {
              bool hit = false;
              if (o[u] < q) {
                // exact rejections of dp_within_col_pf, then the metric
                const double dx = x[0] - y[u][0], dy = x[1] - y[u][1], dz = x[2] - y[u][2];
                double s2 = dx * dx;
                s2 = s2 + dy * dy;
                bool pass;
                if constexpr (NX == 13) {
                  s2 = s2 + dz * dz;
                  pass = !(s2 > T);
                } else {
                  pass = !(s2 > T) && !(S.w[1] * dp_so2(x[2], y[u][2]) > r);
                }
                if (pass) {
                  if constexpr (NX == 13) {
                    double dot = x[3] * y[u][3];
                    dot = dot + x[4] * y[u][4];
                    dot = dot + x[5] * y[u][5];
                    dot = dot + x[6] * y[u][6];
                    hit = !(fabs(dot) < cq) && dp_within_quad_fast(S, x, y[u], r);
                  }
                  else hit = dp_within_planar_fast<NX>(S, x, y[u], r, T);
                }
              }
              const uint64_t mk = __ballot(hit);
              const int pos = cc + __popcll(mk & ltmask);
              if (hit && !pool_pass && pos < slots) conf[(int64_t)q * DP_CONF_SLOTS + pos] = o[u];
              if (hit && pool_pass && pos >= slots) pool[off + (pos - slots)] = o[u];
              cc += __popcll(mk);
            }
            if (lane == c) cnt_l = cc;
          }
        }
      };
      pass(false, lists);
      // lists longer than the inline slots continue in the overflow pool
      const bool over = list_l && cnt_l > slots && cnt_l < DP_CONF_MAX;
      if (over) {
        off_l = (long long)atomicAdd((unsigned long long*)&counters[DPC_CONF], (unsigned long long)(cnt_l - slots));
        if (off_l + cnt_l - slots > pool_cap) cnt_l = DP_CONF_MAX;  // pool exhausted: scan when reached
      }
      const uint64_t pooled = __ballot(over && cnt_l < DP_CONF_MAX);
      if (pooled) {
        const int keep = cnt_l;
        cnt_l = 0;
        pass(true, pooled);
        cnt_l = keep;
        if (over && cnt_l < DP_CONF_MAX) conf_ov[q_l] = (int32_t)off_l;
      }
      if (list_l) {
        conf_cnt[q_l] = (int16_t)(cnt_l < DP_CONF_MAX ? cnt_l : -1);
        // no earlier contender at all: NOVEL unless the cap falls before it
        // (a duplicate of the visited set is not a contender)
        if (cnt_l == 0) flags[q_l] = (int8_t)(flags[q_l] | DP_F_NUC);
      }
    }
  }
}

// List window (planar systems, limit > 0): the contenders of a node after
// its first W in canonical order are flagged DP_F_NOLIST, and their list
// length is set to -(1 + rank) (rank among the node's contenders, at most
// 32766).  The walk reaches one of them only if more than W - limit
// contenders before it were duplicates; it then records 1 + rank in
// need[0], and later batches of the context use W = max(window,
// need + need / 4 + 16) (a fixed window leaves need alone: need == null).
// One wave per node.
__global__ __launch_bounds__(64) void k_list_window(int32_t N, const int32_t* __restrict__ app,
                                                    const int32_t* __restrict__ total,
                                                    const int8_t* __restrict__ verdict, int8_t* __restrict__ flags,
                                                    int16_t* __restrict__ conf_cnt, int32_t window,
                                                    const int32_t* __restrict__ need) {
  const int32_t node = blockIdx.x;
  if (node >= N) return;
  const int lane = threadIdx.x;
  const uint64_t lt = (1ull << lane) - 1ull;
  if (need) {
    const int32_t nd = need[0];
    const int64_t w = (int64_t)nd + nd / 4 + 16;
    if (nd > 0 && w > window) window = w < INT32_MAX ? (int32_t)w : INT32_MAX;
  }
  const int32_t tot = *total;
  const int32_t b = app[node] < tot ? app[node] : tot;
  const int32_t e = app[node + 1] < tot ? app[node + 1] : tot;
  int32_t rank = 0;
  for (int32_t cb = b; cb < e; cb += 64) {
    const int32_t p = cb + lane;
    bool con = false;
    int8_t fl = 0;
    if (p < e && verdict[p] == DP_V_FREE) {
      fl = flags[p];
      con = (fl & 2) || !(fl & 1);
    }
    const uint64_t m = __ballot(con);
    const int32_t r = rank + __popcll(m & lt);
    if (con && r >= window) {
      flags[p] = (int8_t)(fl | DP_F_NOLIST);
      conf_cnt[p] = (int16_t)(-1 - (r < 32766 ? r : 32766));
    }
    rank += __popcll(m);
  }
}

// limit_branching_factor == 0: the reference breaks out of a node's loop
// after its first FREE pair whatever that pair's resolution
// (dbastar.cpp:916-918, size_t >= 0), so every later FREE pair is CAPPED
// before the contender grid is built.  One wave per node.
__global__ __launch_bounds__(64) void k_precap_first(int32_t N, const int32_t* __restrict__ app,
                                                     const int32_t* __restrict__ total,
                                                     const int8_t* __restrict__ verdict, int8_t* __restrict__ flags,
                                                     int8_t* __restrict__ res) {
  const int32_t node = blockIdx.x;
  if (node >= N) return;
  const int lane = threadIdx.x;
  const uint64_t lt = (1ull << lane) - 1ull;
  const int32_t tot = *total;
  const int32_t b = app[node] < tot ? app[node] : tot;
  const int32_t e = app[node + 1] < tot ? app[node + 1] : tot;
  bool seen = false;
  for (int32_t cb = b; cb < e; cb += 64) {
    const int32_t p = cb + lane;
    const bool fr = p < e && verdict[p] == DP_V_FREE;
    const uint64_t m = __ballot(fr);
    if (fr && (seen || (m & lt))) {
      flags[p] = (int8_t)(flags[p] | DP_F_PRECAP);
      res[p] = DP_R_CAPPED;
    }
    seen = seen || m != 0;
  }
}

// ---- Pre-resolution: statuses that are certain before the walk ----------
// A contender without any earlier contender (flag DP_F_NUC, set by
// k_conf_waves) or goal-forced is NOVEL unless its node's cap falls before
// it.  Each conclusion below is a certainty, so the walk reaches the same
// statuses; waiting waves only meet fewer undecided contenders.
//
// One wave per node, its FREE pairs in order: after the limit-th pair that
// is NOVEL unless capped, every FREE pair is CAPPED; a pair that is NOVEL
// unless capped, with fewer than limit possibly NOVEL pairs before it, is
// NOVEL.  The kernel also sets up every node's walk state (the work of
// k_round_init): the cursor at the node's first pair, or at its end when
// every FREE pair was decided here, and the node counted as left to walk
// only in the first case.
__global__ __launch_bounds__(256) void k_pre_nodes(int32_t N, const int32_t* __restrict__ app,
                                                   const int32_t* __restrict__ total,
                                                   const int8_t* __restrict__ verdict,
                                                   const int8_t* __restrict__ flags, int8_t* __restrict__ res,
                                                   int32_t limit, int32_t* __restrict__ cursor,
                                                   int32_t* __restrict__ count, int32_t* __restrict__ node_novel,
                                                   long long* __restrict__ counters) {
  // four nodes per workgroup, one per wave; the counts are summed in LDS and
  // added with one atomic per counter and workgroup (one per node, all at
  // the same moment, queued on the same few addresses)
  __shared__ long long s_add[4];
  if (threadIdx.x < 4) s_add[threadIdx.x] = 0;
  __syncthreads();
  const int32_t node = blockIdx.x * 4 + (int32_t)(threadIdx.x / 64);
  const int lane = threadIdx.x % 64;
  if (node < N) {
    const uint64_t lt = (1ull << lane) - 1ull;
    const int32_t tot = *total;
    const int32_t b = app[node] < tot ? app[node] : tot;
    const int32_t e = app[node + 1] < tot ? app[node + 1] : tot;
    int nuc = 0, possible = 0;
    // statuses this pass leaves the node's FREE pairs in
    int c_nov = 0, c_goal = 0, c_cap = 0;
    bool undec = false;
    // kPre chunks of 64 pairs per step, all their loads issued before the
    // first chunk is evaluated (indices clamped into the node, so the loads
    // are unconditional); past the limit-th such pair the evaluation below
    // caps every later FREE pair
    constexpr int kPre = 4;
    for (int32_t c0 = b; c0 < e; c0 += 64 * kPre) {
      int8_t v_u[kPre], f_u[kPre], r_u[kPre];
#pragma unroll
      for (int u = 0; u < kPre; ++u) {
        const int32_t p = c0 + 64 * u + lane < e ? c0 + 64 * u + lane : e - 1;
        v_u[u] = verdict[p];
        f_u[u] = flags[p];
        r_u[u] = res[p];
      }
#pragma unroll
      for (int u = 0; u < kPre; ++u) {
        const int32_t p = c0 + 64 * u + lane;
        const bool valid = p < e && v_u[u] == DP_V_FREE;
        const int8_t fl = valid ? f_u[u] : 0;
        const int r = valid ? r_u[u] : DP_R_NONE;
        const bool dupv = (fl & 1) && !(fl & 2);
        const bool not_nov = valid && (dupv || r == DP_R_CAPPED);
        const bool nuc_p = valid && (r == DP_R_NOVEL || (fl & (2 | DP_F_NUC)) != 0);
        const uint64_t nuc_m = __ballot(nuc_p);
        const uint64_t pos_m = __ballot(valid && !not_nov);
        const int nuc_before = nuc + __popcll(nuc_m & lt);
        const int pos_before = possible + __popcll(pos_m & lt);
        int st = r;
        if (limit > 0 && nuc_before >= limit) {
          if (valid && r != DP_R_CAPPED) res[p] = DP_R_CAPPED;
          st = DP_R_CAPPED;
        } else if (nuc_p && r != DP_R_NOVEL && (limit <= 0 || pos_before < limit)) {
          res[p] = DP_R_NOVEL;
          st = DP_R_NOVEL;
        }
        if (valid) {
          c_nov += st == DP_R_NOVEL;
          c_goal += st == DP_R_NOVEL && (fl & 2);
          c_cap += st == DP_R_CAPPED;
          undec = undec || (st != DP_R_NOVEL && st != DP_R_CAPPED);
        }
        nuc += __popcll(nuc_m);
        possible += __popcll(pos_m);
      }
    }
    // Every FREE pair decided here: the node is finished as its walk would
    // finish it (its statuses counted, its cursor at the end), and its wave
    // in the walk returns at once.  Otherwise the walk starts at its first
    // pair with nothing counted.
    const bool finished = __ballot(undec) == 0;
    if (finished) {
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) {
        c_nov += __shfl_xor(c_nov, d, 64);
        c_goal += __shfl_xor(c_goal, d, 64);
        c_cap += __shfl_xor(c_cap, d, 64);
      }
    }
    if (lane == 0) {
      cursor[node] = finished ? e : b;
      count[node] = finished ? c_nov : 0;
      if (node_novel) node_novel[node] = finished ? c_nov : 0;
      if (b < e && !finished) atomicAdd((unsigned long long*)&s_add[0], 1ull);
      if (finished) {
        atomicAdd((unsigned long long*)&s_add[1], (unsigned long long)c_nov);
        atomicAdd((unsigned long long*)&s_add[2], (unsigned long long)c_cap);
        atomicAdd((unsigned long long*)&s_add[3], (unsigned long long)c_goal);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_add[0]) atomicAdd((unsigned long long*)&counters[DPC_NODES_LEFT], (unsigned long long)s_add[0]);
    if (s_add[1]) atomicAdd((unsigned long long*)&counters[DPC_NOVEL], (unsigned long long)s_add[1]);
    if (s_add[2]) atomicAdd((unsigned long long*)&counters[DPC_CAPPED], (unsigned long long)s_add[2]);
    if (s_add[3]) atomicAdd((unsigned long long*)&counters[DPC_GOAL], (unsigned long long)s_add[3]);
  }
}

// Counters were zeroed at the start of the batch; nodes with pairs to decide
// are counted into DPC_NODES_LEFT, empty nodes are finished here.
__global__ void k_round_init(int32_t N, const int32_t* __restrict__ app, const int32_t* __restrict__ total,
                             int32_t* __restrict__ cursor, int32_t* __restrict__ count,
                             int32_t* __restrict__ node_novel, long long* counters) {
  const int32_t tot = *total;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const int32_t b = app[i] < tot ? app[i] : tot;
    const int32_t e = app[i + 1] < tot ? app[i + 1] : tot;
    cursor[i] = b;
    count[i] = 0;
    node_novel[i] = 0;
    if (b < e) atomicAdd((unsigned long long*)&counters[DPC_NODES_LEFT], 1ull);
  }
}

struct round_args {
  const dp_dev_sys* sys;
  int32_t N;
  const int32_t* app;
  const int32_t* total;
  const int8_t* verdict;
  const int8_t* flags;
  const int16_t* conf_cnt;
  const int32_t* conf_ov;
  const int32_t* pool;
  const int32_t* conf;
  const double* fin;
  int64_t pcap;
  dp_grid_geom grid;
  const int32_t* cstart;
  const int32_t* cidx;
  const double* cx;
  double r_dup;
  int8_t* res;
  int32_t* cursor;
  int32_t* count;
  int32_t* node_novel;
  int32_t limit;
  int32_t slots;     // inline conflict slots in use (<= DP_CONF_SLOTS)
  int32_t wait;      // polls a lane may spend on one undecided earlier contender
  int32_t poll_cap;  // longest pause between polls, in 64-cycle sleeps
  unsigned long long* trace;  // optional [DP_TRACE_FIELDS * N], fields listed at DP_TRACE_FIELDS
  long long* counters;
  long long* sticky;  // [DPS_COUNT] never reset
  int32_t* wshift;    // quadrotor FP64 window: [0] growth for later batches, [1] the growth this batch used
  int32_t* lneed;     // planar list window: [0] largest 1 + rank of a listless contender a walk reached
};

// Pause between two polls of a contender's verdict: grows with the number of
// polls so far, up to `cap` sleeps of 64 cycles, which keeps thousands of
// waiting waves from flooding the fabric with coherent loads.
__device__ __forceinline__ void poll_pause(int32_t w, int32_t cap) {
  const int32_t k = 1 + (w >> 3) < cap ? 1 + (w >> 3) : cap;
  for (int32_t z = 0; z < k; ++z) __builtin_amdgcn_s_sleep(1);
}

// Advance node `node`'s cursor through its pairs in canonical order until a
// pair waits on an undecided earlier contender of another node.  Returns
// true when the node is finished.  Executed by one whole wave.
// cur / count are wave-uniform registers carried across retries; they are
// also stored for the round kernel, which reloads them in a later launch.
template <int NX>
__device__ bool advance_node(const round_args& a, int32_t node, int32_t& cur, int32_t& count) {
  // per chunk: same-chunk contender masks, undecided-entry masks, flags
  // (1: an earlier contender is NOVEL, 2: an entry past the 64th is
  // undecided), exclusive scan of list lengths, overflow offsets
  __shared__ unsigned long long s_same[64], s_undec[64];
  __shared__ int32_t s_pre[64], s_ov[64], s_flag[64];
  constexpr int kCoop = 16;  // list entries per lane in flight
  const int lane = threadIdx.x & 63;
  const int32_t tot = *a.total;
  const int32_t end = a.app[node + 1] < tot ? a.app[node + 1] : tot;
  if (cur >= end) return true;
  long long n_nov = 0, n_cap = 0, n_dv = 0, n_db = 0, n_goal = 0;
  bool overflow = false;  // a deferred pair (past the FP64 window) was reached before the cap
  // diagnostics (trace only): cycles evaluating lists, cycles in the
  // in-chunk walk, cycles waiting, chunks, waits, polls
  unsigned long long tr_eval = 0, tr_walk = 0, tr_wait = 0, tr_chunks = 0, tr_waits = 0, tr_polls = 0;
  bool stopped = false;
  // per-pair inputs of a chunk (verdict, flags, list length, overflow
  // offset) are loaded together, and those of the next chunk while this one
  // is decided; the values of pairs that are not contenders go unused
  int8_t m_vd = -1, m_fl = 0;
  int16_t m_cc = 0;
  int32_t m_ov = 0;
  if (cur + lane < end) {
    m_vd = a.verdict[cur + lane];
    m_fl = a.flags[cur + lane];
    m_cc = a.conf_cnt[cur + lane];
    m_ov = a.conf_ov[cur + lane];
  }
  for (int32_t cb = cur; cb < end; cb += 64) {
    const int32_t p = cb + lane;
    const int8_t vd_p = m_vd, fl_p = m_fl;
    const int16_t cc_p = m_cc;
    const int32_t ov_p = m_ov;
    m_vd = -1;
    if (p + 64 < end) {
      m_vd = a.verdict[p + 64];
      m_fl = a.flags[p + 64];
      m_cc = a.conf_cnt[p + 64];
      m_ov = a.conf_ov[p + 64];
    }
    // FREE pairs capped before the walk (limit 0) are not walked
    const bool pre = p < end && vd_p == DP_V_FREE && (fl_p & DP_F_PRECAP);
    const bool fr = p < end && vd_p == DP_V_FREE && !pre;
    n_cap += __popcll(__ballot(pre));
    if (a.limit > 0 && count >= a.limit) {
      // The reference has left this node's loop: every later FREE pair is
      // capped.  Other nodes may be waiting for any of them, so the whole
      // tail is published in one pass with its verdict loads in flight
      // together, not chunk by chunk.
      if (fr) __hip_atomic_store(&a.res[p], (int8_t)DP_R_CAPPED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      n_cap += __popcll(__ballot(fr));
      for (int32_t t0 = cb + 64; t0 < end; t0 += 64 * 8) {
        bool f8[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int32_t t = t0 + 64 * k + lane;
          f8[k] = t < end && a.verdict[t] == DP_V_FREE && !(a.flags[t] & DP_F_PRECAP);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          if (f8[k])
            __hip_atomic_store(&a.res[t0 + 64 * k + lane], (int8_t)DP_R_CAPPED, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
          n_cap += __popcll(__ballot(f8[k]));
        }
      }
      break;
    }
    const int8_t fl = fr ? fl_p : 0;
    const bool forced = (fl & 2) != 0;
    const bool dupv = (fl & 1) != 0;
    // past the node's FP64 window: no final state, so it blocks the walk for
    // good; reaching it before the cap is a window overflow
    const bool deferred = (fl & DP_F_DEFER) != 0;
    bool ext_dup = false, ext_block = false;
    uint64_t same = 0;
    // undecided earlier-chunk contenders seen by the first evaluation (the
    // first kBlk; `more` when there were others)
    constexpr int kBlk = 4;
    int32_t blk_o[kBlk];
    int nblk = 0;
    bool more = false;
    uint64_t undec = 0;  // entries (< 64) of a stored list whose contender was undecided
    // entry e of this lane's stored earlier-conflict list
    auto entry = [&](int e) -> int32_t {
      return e < a.slots ? a.conf[(int64_t)p * DP_CONF_SLOTS + e] : a.pool[(int64_t)ov_p + (e - a.slots)];
    };
    // Earlier contenders of pair p: those in this chunk go to `same` (the
    // walk below decides them in order); an earlier-chunk contender already
    // NOVEL makes p a duplicate, one still undecided blocks p.  With
    // polls > 0 an undecided contender is waited for (it belongs to an
    // earlier node, whose wave is resident in persistent mode).
    auto eval = [&](int32_t p, int32_t polls) {
      auto visit = [&](int32_t o) {
        if (o >= cb) {
          same |= 1ull << (o - cb);
          return true;
        }
        int r = polls == 0 ? a.res[o] : __hip_atomic_load(&a.res[o], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int32_t w = 0; r == DP_R_UNDECIDED && w < polls; ++w) {
          poll_pause(w, a.poll_cap);
          r = __hip_atomic_load(&a.res[o], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (r == DP_R_NOVEL) {
          ext_dup = true;
        } else if (r == DP_R_UNDECIDED) {
          ext_block = true;
          if (nblk < kBlk) blk_o[nblk++] = o;
          else more = true;
        }
        return !ext_dup;
      };
      // No stored list: scan the contender grid.  An earlier-chunk
      // contender already decided as not NOVEL cannot matter, so its verdict
      // is read first and the metric skipped for it.
      double x[NX];
#pragma unroll
      for (int k = 0; k < NX; ++k) x[k] = a.fin[(int64_t)k * a.pcap + p];
      scan_earlier<NX>(*a.sys, a.grid, a.cstart, a.cidx, a.cx, a.pcap, x, p, a.r_dup, visit, [&](int32_t o) {
        if (o >= cb) return true;
        const int v = a.res[o];
        return v == DP_R_NOVEL || v == DP_R_UNDECIDED;
      });
    };
    const unsigned long long t_eval = a.trace ? (unsigned long long)wall_clock64() : 0ull;
    const int cnt = fr && !forced && !dupv && !deferred ? cc_p : 0;
    {
      // Stored lists of the chunk's 64 pairs, evaluated by the whole wave:
      // the entries are numbered consecutively over the pairs (exclusive
      // scan of the counts) and lane l takes entries l, l+64, ..., so a long
      // list costs no more than its share.  Plain (cached) loads suffice: a
      // decided verdict never changes, and a stale "undecided" is polled
      // coherently later.
      const int n_l = cnt > 0 ? cnt : 0;
      int incl = n_l;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
      }
      const int total_e = __shfl(incl, 63, 64);
      s_pre[lane] = incl - n_l;
      s_ov[lane] = n_l > a.slots ? ov_p : 0;
      s_same[lane] = 0;
      s_undec[lane] = 0;
      s_flag[lane] = 0;
      __syncthreads();
      for (int f0 = 0; f0 < total_e; f0 += 64 * kCoop) {
        int32_t os[kCoop];
        int rs[kCoop], oe[kCoop];  // oe: owner lane << 16 | entry index, or -1
#pragma unroll
        for (int k = 0; k < kCoop; ++k) {
          const int f = f0 + 64 * k + lane;
          oe[k] = -1;
          os[k] = INT32_MAX;
          if (f < total_e) {
            int lo = 0;
#pragma unroll
            for (int st = 32; st > 0; st >>= 1)  // lo + st <= 63
              if (s_pre[lo + st] <= f) lo += st;
            const int e = f - s_pre[lo];  // < DP_CONF_MAX
            oe[k] = (lo << 16) | e;
            const int32_t q = cb + lo;
            os[k] = e < a.slots ? a.conf[(int64_t)q * DP_CONF_SLOTS + e] : a.pool[(int64_t)s_ov[lo] + (e - a.slots)];
          }
        }
#pragma unroll
        for (int k = 0; k < kCoop; ++k) rs[k] = os[k] < cb ? a.res[os[k]] : (int)DP_R_NONE;
#pragma unroll
        for (int k = 0; k < kCoop; ++k) {
          if (oe[k] < 0) continue;
          const int own = oe[k] >> 16, ee = oe[k] & 0xffff;
          if (os[k] >= cb) {
            atomicOr(&s_same[own], 1ull << (os[k] - cb));
          } else if (rs[k] == DP_R_NOVEL) {
            atomicOr(&s_flag[own], 1);
          } else if (rs[k] == DP_R_UNDECIDED) {
            if (ee < 64) atomicOr(&s_undec[own], 1ull << ee);
            else atomicOr(&s_flag[own], 2);
          }
        }
      }
      __syncthreads();
      if (cnt > 0) {
        same = s_same[lane];
        undec = s_undec[lane];
        ext_dup = (s_flag[lane] & 1) != 0;
        more = (s_flag[lane] & 2) != 0;
        ext_block = !ext_dup && (undec != 0 || more);
      }
      __syncthreads();  // the arrays are rewritten by the next chunk
    }
    if (a.lneed && __ballot(cnt < 0)) {
      // a contender past the list window was reached before the cap: later
      // batches list at least this many contenders per node
      int need = cnt < 0 && (fl & DP_F_NOLIST) ? -cnt : 0;
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) need = max(need, __shfl_xor(need, d, 64));
      if (lane == 0 && need > 0) atomicMax(a.lneed, need);
    }
    if (cnt < 0) eval(p, 0);
    unsigned long long t_walk = 0;
    if (a.trace) {
      t_walk = (unsigned long long)wall_clock64();
      tr_eval += t_walk - t_eval;
      ++tr_chunks;
    }
    const uint64_t free_m = __ballot(fr);
    const uint64_t forced_m = __ballot(fr && forced);
    const uint64_t dupv_m = __ballot(fr && dupv && !forced);
    uint64_t dup_m = __ballot(ext_dup);
    const uint64_t def_m = __ballot(fr && deferred);
    uint64_t block_m = __ballot(ext_block && !ext_dup) | def_m;
    uint64_t novel_m = 0, capped_m = 0, dupb_m = 0, dupvr_m = 0, pub_m = 0;
    int stop_lane = 64;
    // Store the verdicts decided since the last call: other nodes' waves may
    // be waiting for them, so they go out before this wave waits itself.
    const uint64_t lane_bit = 1ull << lane;
    auto publish = [&]() {
      const uint64_t dec = (novel_m | capped_m | dupb_m | dupvr_m) & ~pub_m;
      if (fr && (dec & lane_bit)) {
        const int r = (capped_m & lane_bit) ? DP_R_CAPPED
                      : (novel_m & lane_bit) ? DP_R_NOVEL
                      : (dupvr_m & lane_bit) ? DP_R_DUP_VISITED
                                             : DP_R_DUP_BATCH;
        __hip_atomic_store(&a.res[p], (int8_t)r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      pub_m |= dec;
    };
    // The sequential walk over the chunk, done wave-wide.  Below the first
    // blocked pair fb every status is determined: forced pairs are NOVEL,
    // duplicates of the visited set and pairs with a NOVEL earlier-chunk
    // contender are decided, and the rest are NOVEL unless a same-chunk
    // earlier contender is NOVEL.  That last rule follows the list order, a
    // DAG on the lanes, so iterating it from any start reaches its unique
    // fixed point within (longest chain + 1) rounds.  The branching cap then
    // cuts after the (limit - count)-th NOVEL pair: every later FREE pair is
    // CAPPED, including fb, which then needs no wait.
    bool refreshed = false;
    for (;;) {
      const int fb = block_m ? __ffsll((long long)block_m) - 1 : 64;
      const uint64_t below = fb >= 64 ? ~0ull : ((1ull << fb) - 1ull);
      const uint64_t forced_b = forced_m & below;
      const uint64_t cand = free_m & below & ~forced_m & ~dupv_m & ~dup_m;
      const bool mine = (cand & lane_bit) != 0;
      uint64_t nov = forced_b | cand;
      for (int it = 0; it <= 64; ++it) {
        const uint64_t nx = forced_b | __ballot(mine && (same & nov) == 0);
        if (nx == nov) break;
        nov = nx;
      }
      if (a.limit > 0 && count + __popcll(nov) >= a.limit) {
        // lanes up to the (limit - count)-th NOVEL keep their status
        uint64_t m = nov;
        for (int t = count + 1; t < a.limit; ++t) m &= m - 1ull;
        const int c = __ffsll((long long)m) - 1;
        const uint64_t upto = c >= 63 ? ~0ull : ((2ull << c) - 1ull);
        novel_m = nov & upto;
        dupvr_m = dupv_m & upto;
        dupb_m = free_m & upto & ~novel_m & ~dupvr_m;
        capped_m = free_m & ~upto;
        break;
      }
      novel_m = nov;
      dupvr_m = dupv_m & below;
      dupb_m = free_m & below & ~nov & ~dupv_m;
      if (fb < 64 && ((def_m >> fb) & 1ull)) {
        // Window overflow: the cap does not fall before a pair without FP64
        // data.  Every later FREE pair of the node is published CAPPED so
        // that no wave waits on it; the batch is resolved again afterwards.
        capped_m = free_m & ~below;
        overflow = true;
        break;
      }
      if (fb >= 64) break;
      // Only lane fb's verdict is needed next: wait for its undecided
      // contenders, after publishing what is decided (other nodes' waves
      // may be waiting for it).
      publish();
      if (!refreshed) {
        // First read the undecided contenders of every blocked lane again,
        // all lanes at once: contenders decided since the chunk's list
        // evaluation then cost one round trip instead of one wait each.
        refreshed = true;
        bool nov_found = false, still = false;
        if ((block_m & lane_bit) && cnt >= 0) {
          for (int g = 0; g < 64 && !nov_found; g += 8) {
            if (!((undec >> g) & 0xffull)) continue;
            int32_t oo[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) oo[k] = (undec >> (g + k)) & 1ull ? entry(g + k) : -1;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
              if (oo[k] < 0) continue;
              const int r = __hip_atomic_load(&a.res[oo[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              if (r == DP_R_NOVEL) nov_found = true;
              else if (r != DP_R_UNDECIDED) undec &= ~(1ull << (g + k));
            }
          }
          still = !nov_found && (undec != 0 || more);
        } else if (block_m & lane_bit) {
          still = true;  // grid-scanned lane: its wait re-scans
        }
        still = still || (def_m & lane_bit) != 0;
        dup_m |= __ballot(nov_found);
        const uint64_t nb = __ballot(still);
        if (!(nb & (1ull << fb))) {
          block_m = nb;
          continue;  // fb is decided now: walk on
        }
        block_m = nb;
      }
      int32_t polled = 0;
      const unsigned long long t_wait = a.trace ? (unsigned long long)wall_clock64() : 0ull;
        bool d = false, blk = false;
        if (a.wait > 0 && lane == fb && cnt >= 0) {
          // stored list: poll the undecided contenders, eight at a time,
          // until one is NOVEL or all are decided (entries past the 64th
          // are re-read each round)
          uint64_t m = undec;
          bool mo = more;
          int32_t w = 0;
          for (;;) {
            for (int g = 0; g < 64 && !d; g += 8) {
              if (!((m >> g) & 0xffull)) continue;
              int32_t oo[8];
#pragma unroll
              for (int k = 0; k < 8; ++k) oo[k] = (m >> (g + k)) & 1ull ? entry(g + k) : -1;
#pragma unroll
              for (int k = 0; k < 8; ++k) {
                const int r = oo[k] >= 0
                                  ? __hip_atomic_load(&a.res[oo[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : (int)DP_R_UNDECIDED;
                if (oo[k] < 0) continue;
                if (r == DP_R_NOVEL) d = true;
                else if (r != DP_R_UNDECIDED) m &= ~(1ull << (g + k));
              }
            }
            if (mo && !d) {
              mo = false;
              for (int g = 64; g < cnt && !d; g += 8) {
                int32_t oo[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) oo[k] = g + k < cnt ? entry(g + k) : INT32_MAX;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                  if (oo[k] >= cb) continue;  // past the list, or in this chunk
                  const int r = __hip_atomic_load(&a.res[oo[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                  if (r == DP_R_NOVEL) d = true;
                  else if (r == DP_R_UNDECIDED) mo = true;
                }
              }
            }
            if (d || (!m && !mo)) break;
            if (++w > a.wait) {
              blk = true;
              break;
            }
            poll_pause(w, a.poll_cap);
          }
          polled = w;
        } else if (a.wait > 0 && lane == fb) {
          if (!more) {
            // wait for the recorded contenders only
            for (int e = 0; e < nblk && !d && !blk; ++e) {
              int r = __hip_atomic_load(&a.res[blk_o[e]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              int32_t w = 0;
              for (; r == DP_R_UNDECIDED && w < a.wait; ++w) {
                poll_pause(w, a.poll_cap);
                r = __hip_atomic_load(&a.res[blk_o[e]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              }

              if (r == DP_R_NOVEL) d = true;
              else if (r == DP_R_UNDECIDED) blk = true;
            }
          } else {
            ext_dup = false;
            ext_block = false;
            nblk = 0;
            eval(p, a.wait);
            d = ext_dup;
            blk = ext_block && !ext_dup;
          }
        }
      if (a.trace) {
        const unsigned long long t1 = (unsigned long long)wall_clock64();
        tr_wait += t1 - t_wait;
        t_walk += t1 - t_wait;  // the walk clock excludes the wait
        ++tr_waits;
        tr_polls += (unsigned long long)__shfl(polled, fb, 64);
      }
      if (a.wait <= 0 || __ballot(blk) != 0) {
        stop_lane = fb;
        break;
      }
      block_m &= ~(1ull << fb);
      if (__ballot(d) != 0) dup_m |= 1ull << fb;
      refreshed = false;  // later blocked lanes may have been decided meanwhile
    }
    publish();
    if (a.trace) tr_walk += (unsigned long long)wall_clock64() - t_walk;
    count += __popcll(novel_m);
    n_nov += __popcll(novel_m);
    n_goal += __popcll(novel_m & forced_m);  // goal-forced pairs the reference reached
    n_cap += __popcll(capped_m);
    n_dv += __popcll(dupvr_m);
    n_db += __popcll(dupb_m);
    if (stop_lane < 64) {
      cur = cb + stop_lane;
      stopped = true;
      break;
    }
    if (overflow) {
      // the rest of the node's FREE pairs go out CAPPED, as at the cap
      count = a.limit > count ? a.limit : count;
      if (lane == 0) {
        atomicAdd((unsigned long long*)&a.counters[DPC_WINDOW_OVF], 1ull);
        if (a.sticky) atomicAdd((unsigned long long*)&a.sticky[DPS_WINDOW], 1ull);
        // later batches of this context use a window twice as wide
        if (a.wshift) atomicMax(&a.wshift[0], a.wshift[1] + 1);
      }
      overflow = false;
    }
  }
  if (!stopped) cur = end;
  if (a.trace && lane == 0) {
    unsigned long long* t = a.trace + (int64_t)DP_TRACE_FIELDS * node;
    t[2] += tr_wait;
    t[3] += tr_eval;
    t[4] += tr_walk;
    t[5] += tr_chunks;
    t[6] += tr_waits;
    t[7] += tr_polls;
  }
  if (lane == 0) {
    a.cursor[node] = cur;
    a.count[node] = count;
    if (!stopped) {
      if (a.node_novel) a.node_novel[node] = count;
      atomicAdd((unsigned long long*)&a.counters[DPC_NODES_LEFT], (unsigned long long)-1ll);
    }
    if (n_nov) atomicAdd((unsigned long long*)&a.counters[DPC_NOVEL], (unsigned long long)n_nov);
    if (n_cap) atomicAdd((unsigned long long*)&a.counters[DPC_CAPPED], (unsigned long long)n_cap);
    if (n_dv) atomicAdd((unsigned long long*)&a.counters[DPC_DUPV], (unsigned long long)n_dv);
    if (n_db) atomicAdd((unsigned long long*)&a.counters[DPC_DUPB], (unsigned long long)n_db);
    if (n_goal) atomicAdd((unsigned long long*)&a.counters[DPC_GOAL], (unsigned long long)n_goal);
  }
  return !stopped;
}

// One resolution round (host-driven completion path): one wave per node.
template <int NX>
__global__ __launch_bounds__(64) void k_round(round_args a) {
  if (__hip_atomic_load(&a.counters[DPC_NODES_LEFT], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
  const int32_t node = blockIdx.x;
  if (node >= a.N) return;
  if (node == 0 && threadIdx.x == 0) atomicAdd((unsigned long long*)&a.counters[DPC_ROUNDS], 1ull);
  int32_t cur = a.cursor[node], count = a.count[node];
  (void)advance_node<NX>(a, node, cur, count);
}

// Persistent resolution: waves claim nodes in ascending order through a
// ticket and advance them until finished, backing off while blocked.  A
// wave only waits for nodes with smaller tickets, all held by resident
// waves, so the smallest unfinished node never waits.  The spin is bounded:
// a wave that exhausts it leaves its node to the round kernel and raises a
// flag the host checks.
template <int NX>
__device__ void walk_node(const round_args& a, int32_t node) {
  if (a.trace && threadIdx.x == 0) {
    unsigned long long* t = a.trace + (int64_t)DP_TRACE_FIELDS * node;
    t[0] = (unsigned long long)wall_clock64();
    for (int k = 2; k < DP_TRACE_FIELDS; ++k) t[k] = 0;
  }
  __syncthreads();
  int32_t spins = 0;
  int32_t cur = a.cursor[node], count = a.count[node];
  const int32_t limit = a.wait > 0 ? DP_SPIN_LIMIT_WAITING : DP_SPIN_LIMIT;
  while (!advance_node<NX>(a, node, cur, count)) {
    if (++spins > limit) {
      if (threadIdx.x == 0) {
        atomicOr((unsigned long long*)&a.counters[DPC_ERRFLAG], 2ull);
        if (a.sticky) atomicAdd((unsigned long long*)&a.sticky[DPS_GIVEUP], 1ull);
      }
      break;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  if (threadIdx.x == 0 && spins) atomicAdd((unsigned long long*)&a.counters[DPC_ROUNDS], (unsigned long long)spins);
  if (a.trace && threadIdx.x == 0) a.trace[(int64_t)DP_TRACE_FIELDS * node + 1] = (unsigned long long)wall_clock64();
}

template <int NX>
__global__ __launch_bounds__(64) void k_resolve_ticketed(round_args a) {
  __shared__ int32_t s_node;
  if (threadIdx.x == 0) s_node = (int32_t)atomicAdd((unsigned long long*)&a.counters[DPC_TICKET], 1ull);
  __syncthreads();
  const int32_t node = s_node;
  if (node >= a.N) return;
  walk_node<NX>(a, node);
}

// The ordered walks of many instances of one system in one launch.  Waves
// claim (instance, node) items from one queue in ticket order; the host
// interleaves the instances in short runs of nodes, so every instance's
// chain of dependent nodes advances at the same time instead of one
// instance after another.  Within an instance the tickets ascend with the
// node, so the argument of k_resolve_ticketed holds per instance: a wave
// only waits for nodes of its instance with smaller tickets, all claimed
// earlier by resident or finished waves.  The wave that finishes an
// instance's last node stamps the instance's completion time.
template <int NX>
__global__ __launch_bounds__(64) void k_resolve_queue(const round_args* __restrict__ args,
                                                      const int2* __restrict__ items, int32_t n_items,
                                                      unsigned long long* __restrict__ queue,
                                                      int32_t* __restrict__ done,
                                                      unsigned long long* __restrict__ stamps) {
  __shared__ int32_t s_t;
  if (threadIdx.x == 0) s_t = (int32_t)atomicAdd(queue, 1ull);
  __syncthreads();
  const int32_t t = __builtin_amdgcn_readfirstlane(s_t);
  if (t >= n_items) return;
  const int2 it = items[t];
  const int32_t inst = __builtin_amdgcn_readfirstlane(it.x);
  const int32_t node = __builtin_amdgcn_readfirstlane(it.y);
  const round_args& a = args[inst];
  walk_node<NX>(a, node);
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&done[inst], 1) + 1 == a.N) stamps[inst] = (unsigned long long)wall_clock64();
  }
}

__global__ void k_queue_reset(unsigned long long* __restrict__ queues, int32_t nq, int32_t* __restrict__ done,
                              int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nq) queues[i] = 0ull;
  if (i < n) done[i] = 0;
}

// Append the NOVEL final states of the batch to the visited set in canonical
// order; dst holds the exclusive scan of NOVEL pairs per node.
template <int NX>
__global__ __launch_bounds__(64) void k_gather_novel(int32_t N, const int32_t* __restrict__ app,
                                                     const int32_t* __restrict__ total,
                                                     const int8_t* __restrict__ res, const int32_t* __restrict__ dst,
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

template <typename F>
int dispatch_nx(int nx, F&& f) {
  switch (nx) {
    case 3: return f(std::integral_constant<int, 3>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 5: return f(std::integral_constant<int, 5>{});
    case 13: return f(std::integral_constant<int, 13>{});
  }
  dp_set_error("unsupported nx");
  return DP_ERR_ARG;
}

}  // namespace

static round_args make_round_args(dp_ctx* c) {
  round_args a{};
  a.sys = c->d_sys;
  a.N = c->N;
  a.app = c->d_app;
  a.total = c->d_total;
  a.verdict = c->d_verdict;
  a.flags = c->d_flags;
  a.conf_cnt = c->d_conf_cnt;
  a.conf_ov = c->d_conf_ov;
  a.pool = c->d_conf_pool;
  a.conf = c->d_conf;
  a.fin = c->d_final;
  a.pcap = c->pair_cap;
  a.grid = c->grid;
  a.cstart = c->d_cont_start;
  a.cidx = c->d_cont_idx;
  a.cx = c->d_cont_x;
  a.r_dup = c->round_r_dup;
  a.res = c->d_res;
  a.cursor = c->d_node_cursor;
  a.count = c->d_node_count;
  a.node_novel = c->d_node_novel;
  a.limit = c->round_limit > 0 ? c->round_limit : 0;  // 0 (no cap here) also for limit 0: k_precap_first
  a.slots = c->conf_slots;
  a.wait = 0;
  a.poll_cap = c->poll_cap;
  a.trace = c->d_trace;
  a.counters = c->d_counters;
  a.sticky = c->d_sticky;
  a.wshift = c->d_wshift;
  a.lneed = c->fp64_window > 0 ? nullptr : c->d_lneed;
  return a;
}

int dp_launch_visited_rebuild(dp_ctx* c, double radius, hipStream_t s) {
  c->grid = dp_make_grid(c, radius);
  c->grid_radius = radius;
  const int64_t nc = c->grid.ncells;
  int rc = dp_ensure_i32(&c->d_vis_start, &c->vis_cells_cap, nc + 1);
  if (rc) return rc;
  dp_fill_i32(c->d_vis_start, nc + 1, 0, s);
  if (c->nv > 0) {
    rc = dp_ensure_i32(&c->d_cont_cursor, &c->cont_cursor_cap, nc + 1);
    if (rc) return rc;
    rc = dispatch_nx(c->desc.nx, [&](auto NXc) -> int {
      constexpr int NX = decltype(NXc)::value;
      k_vis_count<NX><<<dp_grid_blocks(c->nv, 256, 8192), 256, 0, s>>>(c->grid, c->d_vis, c->nv_cap, c->nv,
                                                                      c->d_vis_cell, c->d_vis_start);
      (void)dp_scan_large(c, c->d_vis_start, nullptr, nc, c->d_counters, -1, s);
      dp_fill_i32(c->d_cont_cursor, nc, 0, s);
      k_vis_scatter<NX><<<dp_grid_blocks(c->nv, 256, 8192), 256, 0, s>>>(c->d_vis, c->nv_cap, c->nv, c->d_vis_cell,
                                                                        c->d_vis_start, c->d_cont_cursor,
                                                                        c->d_vis_sorted, c->d_vis_perm);
      return DP_OK;
    });
    if (rc) return rc;
  }
  DP_HIP_RET(hipGetLastError());
  c->grid_dirty = false;
  dp_bump_epoch();
  return DP_OK;
}

int dp_launch_resolve(dp_ctx* c, const dp_expand_params* p, double r_dup, hipStream_t s) {
  const int64_t nc = c->grid.ncells;
  const int32_t N = c->N;
  // the scatter cursors, then the conflict-wave items: at most one per
  // contender
  int rc = dp_ensure_i32(&c->d_cont_start, &c->cont_start_cap, nc + 1);
  if (!rc) rc = dp_ensure_i32(&c->d_cont_cursor, &c->cont_cursor_cap, std::max<int64_t>(nc + 1, c->pair_cap));
  if (rc) return rc;
  c->round_limit = p->limit_branching_factor;
  c->round_r_dup = r_dup;
  c->round_params = *p;
  // Planar systems with a cap: only the first DP_FP64_WINDOW(limit) FREE
  // pairs of each node get a stored list (or the test limit).  Quadrotor
  // pairs past their node's FP64 window are not contenders at all
  // (k_quad_window), so every quadrotor contender gets a list.
  const int32_t lim = p->limit_branching_factor;
  const int32_t window = lim > 0 && c->desc.system != DP_SYS_QUADROTOR && !c->select_mode
                             ? (c->fp64_window > 0 ? c->fp64_window : DP_FP64_WINDOW(lim))
                             : 0;
  const int32_t* lneed = c->fp64_window > 0 ? nullptr : c->d_lneed;
  const int pb = dp_grid_blocks(c->pair_cap, 256, 8192);
  const int64_t pool_cap = c->pool_limit > 0 ? std::min<int64_t>(c->pool_limit, c->conf_pool_cap) : c->conf_pool_cap;
  rc = dispatch_nx(c->desc.nx, [&](auto NXc) -> int {
    constexpr int NX = decltype(NXc)::value;
    if (p->limit_branching_factor == 0 && N > 0)
      k_precap_first<<<N, 64, 0, s>>>(N, c->d_app, c->d_total, c->d_verdict, c->d_flags, c->d_res);
    if (window > 0 && N > 0)
      k_list_window<<<N, 64, 0, s>>>(N, c->d_app, c->d_total, c->d_verdict, c->d_flags, c->d_conf_cnt, window,
                                     lneed);
    dp_fill_i32(c->d_cont_start, nc + 1, 0, s);
    k_cont_count<NX><<<pb, 256, 0, s>>>(c->grid, c->d_total, c->d_verdict, c->d_flags, c->d_final, c->pair_cap,
                                        c->d_cont_start);
    (void)dp_scan_large(c, c->d_cont_start, nullptr, nc, c->d_counters, -1, s);
    dp_fill_i32(c->d_cont_cursor, nc, 0, s);
    k_cont_scatter<NX><<<pb, 256, 0, s>>>(c->grid, c->d_total, c->d_verdict, c->d_flags, c->d_final, c->pair_cap,
                                          c->d_cont_start, c->d_cont_cursor, c->d_cont_idx, c->d_cont_x);
    // earlier-conflict lists; the scatter cursors are free again and hold
    // the list of non-empty cells
    k_cell_list<<<(int)((nc + 255) / 256), 256, 0, s>>>(nc, c->grid, c->d_cont_start, c->d_cont_idx, window,
                                                                64 * (NX == 13 ? 2 : 4), c->d_flags, c->d_conf_cnt,
                                                                c->d_cont_cursor, c->d_counters);
    k_conf_waves<NX><<<(int)std::min<int64_t>((nc + 3) / 4, 8192), 256, 0, s>>>(
        c->d_sys, c->grid, c->d_flags, c->d_pnode, c->d_app, window, c->pair_cap, c->d_cont_start, c->d_cont_idx, c->d_cont_x, c->d_cont_cursor,
        r_dup, c->d_conf_cnt, c->d_conf, c->d_conf_ov, c->d_conf_pool, pool_cap, c->conf_slots, c->d_counters);
    if (dp_record_event(c, 7, s)) return DP_ERR_HIP;
    if (p->limit_branching_factor > 0)
      k_pre_nodes<<<(N + 3) / 4, 256, 0, s>>>(N, c->d_app, c->d_total, c->d_verdict, c->d_flags, c->d_res,
                                               p->limit_branching_factor, c->d_node_cursor, c->d_node_count,
                                               c->d_node_novel, c->d_counters);
    else
      k_round_init<<<dp_grid_blocks(N, 256, 1024), 256, 0, s>>>(N, c->d_app, c->d_total, c->d_node_cursor,
                                                                c->d_node_count, c->d_node_novel, c->d_counters);
    if (c->defer_walk) return DP_OK;  // a multi-instance batch walks every instance in one queue
    round_args ra = make_round_args(c);
    ra.wait = DP_WAIT_POLLS;
    k_resolve_ticketed<NX><<<N, 64, 0, s>>>(ra);
    return DP_OK;
  });
  if (rc) return rc;
  DP_HIP_RET(hipGetLastError());
  return DP_OK;
}

namespace {
// Resolution state back to "every walked FREE pair undecided" (pairs capped
// before the walk stay CAPPED), with the resolution counters cleared.
__global__ void k_reset_resolution(const int32_t* __restrict__ total, const int8_t* __restrict__ verdict,
                                   int8_t* __restrict__ flags, int8_t* __restrict__ res, long long* __restrict__ counters) {
  const int32_t tot = *total;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
    if (verdict[q] != DP_V_FREE || (flags[q] & DP_F_PRECAP)) continue;
    res[q] = DP_R_UNDECIDED;
    flags[q] = (int8_t)(flags[q] & ~DP_F_NUC);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int slots[] = {DPC_NOVEL, DPC_DUPV, DPC_DUPB, DPC_CAPPED, DPC_GOAL, DPC_NODES_LEFT, DPC_CONF,
                         DPC_TICKET, DPC_CELLS, DPC_WINDOW_OVF, DPC_ROUNDS};
    for (int k : slots) counters[k] = 0;
    counters[DPC_ERRFLAG] &= ~2ll;
  }
}
}  // namespace

int dp_launch_resolve_again(dp_ctx* c, hipStream_t s) {
  k_reset_resolution<<<dp_grid_blocks(c->pair_cap, 256, 8192), 256, 0, s>>>(c->d_total, c->d_verdict, c->d_flags,
                                                                           c->d_res, c->d_counters);
  DP_HIP_RET(hipGetLastError());
  const dp_expand_params p = c->round_params;
  return dp_launch_resolve(c, &p, c->round_r_dup, s);
}

int dp_launch_more_rounds(dp_ctx* c, int rounds, hipStream_t s) {
  round_args a = make_round_args(c);
  int rc = dispatch_nx(c->desc.nx, [&](auto NXc) -> int {
    constexpr int NX = decltype(NXc)::value;
    for (int r = 0; r < rounds; ++r) k_round<NX><<<c->N, 64, 0, s>>>(a);
    return DP_OK;
  });
  if (rc) return rc;
  DP_HIP_RET(hipGetLastError());
  return DP_OK;
}

// ---- the ordered walks of a multi-instance batch as one work queue ----

struct dp_walk_queue {
  struct klass {
    int nx;
    int32_t first, count;  // its items: d_items[first, first + count)
  };
  int32_t n = 0;  // instances
  round_args* d_args = nullptr;
  int2* d_items = nullptr;
  int64_t items_cap = 0;
  unsigned long long* d_queues = nullptr;  // one ticket per class
  int32_t* d_done = nullptr;               // finished nodes per instance
  std::vector<round_args> h_args;
  std::vector<int2> h_items;
  std::vector<klass> classes;
  bool uploaded = false;
};

namespace {
constexpr int kWalkClasses = 4;  // nx 3, 4, 5, 13
constexpr int32_t kWalkRun = 8;  // nodes of one instance per run of the interleave
}  // namespace

dp_walk_queue* dp_walk_queue_create(int32_t n) {
  if (n <= 0) return nullptr;
  dp_walk_queue* q = new dp_walk_queue;
  q->n = n;
  bool ok = hipMalloc(&q->d_args, sizeof(round_args) * n) == hipSuccess;
  ok = ok && hipMalloc(&q->d_queues, sizeof(unsigned long long) * kWalkClasses) == hipSuccess;
  ok = ok && hipMalloc(&q->d_done, sizeof(int32_t) * n) == hipSuccess;
  if (!ok) {
    dp_walk_queue_destroy(q);
    return nullptr;
  }
  return q;
}

void dp_walk_queue_destroy(dp_walk_queue* q) {
  if (!q) return;
  if (q->d_args) (void)hipFree(q->d_args);
  if (q->d_items) (void)hipFree(q->d_items);
  if (q->d_queues) (void)hipFree(q->d_queues);
  if (q->d_done) (void)hipFree(q->d_done);
  delete q;
}

int dp_walk_queue_prepare(dp_walk_queue* q, dp_ctx* const* ctx, hipStream_t prev) {
  const int32_t n = q->n;
  std::vector<round_args> args(n);
  for (int32_t i = 0; i < n; ++i) {
    std::memset(&args[i], 0, sizeof(round_args));
    args[i] = make_round_args(ctx[i]);
    args[i].wait = DP_WAIT_POLLS;
  }
  std::vector<int2> items;
  std::vector<dp_walk_queue::klass> classes;
  for (int nx : {3, 4, 5, 13}) {
    std::vector<int32_t> mine;
    int32_t nmax = 0;
    for (int32_t i = 0; i < n; ++i)
      if (ctx[i]->desc.nx == nx) {
        mine.push_back(i);
        nmax = std::max(nmax, ctx[i]->N);
      }
    if (mine.empty()) continue;
    dp_walk_queue::klass k{nx, (int32_t)items.size(), 0};
    for (int32_t b = 0; b < nmax; b += kWalkRun)
      for (int32_t i : mine)
        for (int32_t v = b; v < std::min(b + kWalkRun, ctx[i]->N); ++v) items.push_back(make_int2(i, v));
    k.count = (int32_t)items.size() - k.first;
    classes.push_back(k);
  }
  const bool same = q->uploaded && items.size() == q->h_items.size() &&
                    std::memcmp(items.data(), q->h_items.data(), sizeof(int2) * items.size()) == 0 &&
                    std::memcmp(args.data(), q->h_args.data(), sizeof(round_args) * n) == 0;
  if (same) return DP_OK;
  // the previous batch may still read the arrays
  if (prev) DP_HIP_RET(hipStreamSynchronize(prev));
  if ((int64_t)items.size() > q->items_cap) {
    if (q->d_items) DP_HIP_RET(hipFree(q->d_items));
    q->d_items = nullptr;
    q->items_cap = 0;
    DP_HIP_RET(hipMalloc(&q->d_items, sizeof(int2) * items.size()));
    q->items_cap = (int64_t)items.size();
    dp_bump_epoch();
  }
  DP_HIP_RET(hipMemcpy(q->d_args, args.data(), sizeof(round_args) * n, hipMemcpyHostToDevice));
  if (!items.empty())
    DP_HIP_RET(hipMemcpy(q->d_items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice));
  q->h_args.swap(args);
  q->h_items.swap(items);
  q->classes.swap(classes);
  q->uploaded = true;
  return DP_OK;
}

int dp_walk_queue_launch(dp_walk_queue* q, unsigned long long* stamps, hipStream_t s) {
  if (!q->uploaded) {
    dp_set_error("walk queue launched before it was prepared");
    return DP_ERR_STATE;
  }
  const int32_t m = std::max<int32_t>(q->n, kWalkClasses);
  k_queue_reset<<<(m + 255) / 256, 256, 0, s>>>(q->d_queues, kWalkClasses, q->d_done, q->n);
  for (size_t k = 0; k < q->classes.size(); ++k) {
    const dp_walk_queue::klass& K = q->classes[k];
    int rc = dispatch_nx(K.nx, [&](auto NXc) -> int {
      constexpr int NX = decltype(NXc)::value;
      k_resolve_queue<NX><<<K.count, 64, 0, s>>>(q->d_args, q->d_items + K.first, K.count, q->d_queues + k,
                                                 q->d_done, stamps);
      return DP_OK;
    });
    if (rc) return rc;
  }
  DP_HIP_RET(hipGetLastError());
  return DP_OK;
}

int dp_launch_visited_append_novel(dp_ctx* c, int64_t novel_total, hipStream_t s) {
  // Caller guarantees visited capacity >= nv + novel_total.
  const int32_t N = c->N;
  if (N <= 0 || novel_total <= 0) return DP_OK;
  dp_scan_small(c->d_node_novel, N, nullptr, INT64_MAX, c->d_counters, -1, s);
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
  dp_bump_epoch();
  return DP_OK;
}
