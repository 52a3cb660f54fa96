// k_radius.hip -- radius query over the visited set (T_n->nearestR,
// include/dynoplan/nigh_custom_spaces.hpp:236-247; used for rewiring at
// src/dbastar/dbastar.cpp:883-912) on the device grid of the duplicate test.
//
// For each query state: the indices (append order of the visited set) of
// every visited state within `radius` under the system metric, ascending,
// as CSR.  Three kernels:
//   k_radius<NX, false>  one wave per query over the 9 x-contiguous runs of
//                        the query's 27-cell block: the hit count
//   k_radius_offsets     one workgroup: exclusive scan of the counts into
//                        the CSR offsets, plus total and longest list
//   k_radius<NX, true>   the same scan again, hits written at their offsets
//   k_radius_sort        one workgroup per query: its list sorted in LDS
//
// Two entry points share the kernels: dp_visited_radius (host buffers;
// device buffers grown on demand) and dp_visited_radius_noalloc (device
// buffers; no allocation: its per-query counts and status words live in the
// file-scope array `scratch` below, one copy per device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>

#include "dp_internal.h"

// Longest neighbour list the sort kernel orders in LDS (int32 keys).
#define DP_RADIUS_SORT_MAX 4096

namespace {

template <typename F>
int radius_dispatch(int nx, F&& f) {
  switch (nx) {
    case 3: return f(std::integral_constant<int, 3>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 5: return f(std::integral_constant<int, 5>{});
    case 13: return f(std::integral_constant<int, 13>{});
  }
  dp_set_error("unsupported nx");
  return DP_ERR_ARG;
}

struct radius_args {
  const dp_dev_sys* sys;
  dp_grid_geom grid;
  const int32_t* vstart;   // [ncells+1] visited grid cell offsets
  const double* vis;       // [nx][vstride] visited states sorted by cell
  int64_t vstride;
  const int32_t* vperm;    // [nv] sorted position -> append index
  const double* queries;   // [n][nx] row-major
  int64_t n;
  double r;
  int32_t* counts;         // [n]
  const int64_t* offsets;  // [n+1] (fill pass)
  int64_t* idx;            // [idx_cap] (fill pass)
  int64_t idx_cap;
};

// One wave per query.  Lanes 0..8 locate the 9 runs of the 27-cell block
// (rows c1-1..c1+1, layers c2-1..c2+1, x from c0-1 to c0+1 clipped: the rules
// of dp_cell_neighbor), then the candidates of all runs are numbered 0..K-1
// and lane l tests l, l+64, ...; hits are ballot-compacted.
template <int NX, bool FILL>
__global__ __launch_bounds__(256) void k_radius(radius_args a) {
  const dp_dev_sys& S = *a.sys;
  const dp_grid_geom& G = a.grid;
  const int lane = threadIdx.x % 64;
  const uint64_t lt = (1ull << lane) - 1ull;
  const double T = dp_pos_bound(S, a.r);
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / 64);
  for (int64_t q = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; q < a.n; q += nwaves) {
    double x[NX];
#pragma unroll
    for (int k = 0; k < NX; ++k) x[k] = a.queries[q * NX + k];
    int32_t c[3];
    dp_cell_coords(G, x, c);
    int len = 0, st = 0;
    if (lane < 9) {
      const int32_t n1 = c[1] + lane % 3 - 1;
      const int32_t d2 = lane / 3 - 1;
      int32_t n2 = c[2] + d2;
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
        const int32_t x0 = c[0] > 0 ? c[0] - 1 : 0, x1 = c[0] + 1 < G.dims[0] ? c[0] + 1 : G.dims[0] - 1;
        st = a.vstart[row + x0];
        len = a.vstart[row + x1 + 1] - st;
      }
    }
    int incl = len;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
      const int v = __shfl_up(incl, d, 64);
      if (lane >= d) incl += v;
    }
    int32_t rs[9], rp[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      rs[k] = __builtin_amdgcn_readlane(st, k);
      rp[k] = __builtin_amdgcn_readlane(incl - len, k);
    }
    const int K = __builtin_amdgcn_readlane(incl, 8);
    const int64_t base = FILL ? a.offsets[q] : 0;
    int64_t cnt = 0;
    for (int j0 = 0; j0 < K; j0 += 64) {
      const int j = j0 + lane;
      bool hit = false;
      int32_t t = 0;
      if (j < K) {
        int k = 0;
#pragma unroll
        for (int kk = 1; kk < 9; ++kk) k += j >= rp[kk] ? 1 : 0;
        int32_t sk = rs[0], pk = rp[0];
#pragma unroll
        for (int kk = 1; kk < 9; ++kk)
          if (k == kk) sk = rs[kk], pk = rp[kk];
        t = sk + (j - pk);
        hit = dp_within_col_pf<NX>(S, x, a.vis, a.vstride, t, a.r, T);
      }
      const uint64_t m = __ballot(hit);
      if (FILL && hit) {
        const int64_t pos = base + cnt + __popcll(m & lt);
        if (pos < a.idx_cap) a.idx[pos] = a.vperm[t];
      }
      cnt += __popcll(m);
    }
    if (!FILL && lane == 0) a.counts[q] = (int32_t)cnt;
  }
}

// Exclusive scan of counts[0..n) into offsets[0..n], one workgroup of 1024
// threads, each over a contiguous chunk.  status[0] = total, status[1] =
// longest list.
__global__ __launch_bounds__(1024) void k_radius_offsets(const int32_t* __restrict__ counts, int64_t n,
                                                         int64_t* __restrict__ offsets, long long* __restrict__ status) {
  __shared__ long long s[1024];
  __shared__ int smax[1024];
  const int t = threadIdx.x;
  const int64_t chunk = (n + 1023) / 1024;
  const int64_t b = t * chunk < n ? t * chunk : n;
  const int64_t e = b + chunk < n ? b + chunk : n;
  long long sum = 0;
  int mx = 0;
  for (int64_t k = b; k < e; ++k) {
    sum += counts[k];
    mx = counts[k] > mx ? counts[k] : mx;
  }
  s[t] = sum;
  smax[t] = mx;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const long long u = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += u;
    __syncthreads();
  }
  for (int off = 512; off > 0; off >>= 1) {
    if (t < off) smax[t] = smax[t + off] > smax[t] ? smax[t + off] : smax[t];
    __syncthreads();
  }
  long long run = s[t] - sum;
  for (int64_t k = b; k < e; ++k) {
    offsets[k] = run;
    run += counts[k];
  }
  if (t == 1023) {
    offsets[n] = s[1023];
    status[0] = s[1023];
    status[1] = smax[0];
  }
}

// One workgroup per query: its list (when it lies below idx_cap and holds
// at most DP_RADIUS_SORT_MAX entries) sorted ascending by a bitonic network
// in LDS, padded with INT32_MAX to a power of two.
__global__ __launch_bounds__(256) void k_radius_sort(const int64_t* __restrict__ offsets, int64_t n,
                                                     int64_t* __restrict__ idx, int64_t idx_cap) {
  __shared__ int32_t s[DP_RADIUS_SORT_MAX];
  for (int64_t q = blockIdx.x; q < n; q += gridDim.x) {
    const int64_t b = offsets[q], e = offsets[q + 1];
    const int64_t len = e - b;
    if (len < 2 || len > DP_RADIUS_SORT_MAX || e > idx_cap) continue;
    int p2 = 2;
    while (p2 < len) p2 <<= 1;
    for (int i = threadIdx.x; i < p2; i += blockDim.x) s[i] = i < len ? (int32_t)idx[b + i] : INT32_MAX;
    __syncthreads();
    for (int k = 2; k <= p2; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < p2; i += blockDim.x) {
          const int ij = i ^ j;
          if (ij > i) {
            const int32_t u = s[i], v = s[ij];
            const bool up = (i & k) == 0;
            if ((u > v) == up) {
              s[i] = v;
              s[ij] = u;
            }
          }
        }
        __syncthreads();
      }
    }
    for (int i = threadIdx.x; i < len; i += blockDim.x) idx[b + i] = s[i];
    __syncthreads();
  }
}

int launch_radius(const dp_ctx* c, const double* d_queries, int64_t n, double radius, int32_t* d_counts,
                  int64_t* d_offsets, int64_t* d_idx, int64_t idx_cap, long long* d_status, bool fill_only,
                  hipStream_t s) {
  radius_args a{};
  a.sys = c->d_sys;
  a.grid = c->grid;
  a.vstart = c->d_vis_start;
  a.vis = c->d_vis_sorted;
  a.vstride = c->nv_cap;
  a.vperm = c->d_vis_perm;
  a.queries = d_queries;
  a.n = n;
  a.r = radius;
  a.counts = d_counts;
  a.offsets = d_offsets;
  a.idx = d_idx;
  a.idx_cap = idx_cap;
  const int blocks = dp_grid_blocks(n, 4, 8192);
  return radius_dispatch(c->desc.nx, [&](auto NXc) -> int {
    constexpr int NX = decltype(NXc)::value;
    if (!fill_only) {
      k_radius<NX, false><<<blocks, 256, 0, s>>>(a);
      k_radius_offsets<<<1, 1024, 0, s>>>(d_counts, n, d_offsets, d_status);
    } else {
      k_radius<NX, true><<<blocks, 256, 0, s>>>(a);
      k_radius_sort<<<dp_grid_blocks(n, 1, 8192), 256, 0, s>>>(d_offsets, n, d_idx, idx_cap);
    }
    DP_HIP_RET(hipGetLastError());
    return DP_OK;
  });
}

// ---- device buffers of the allocating entry point ---------------------------
struct radius_bufs {
  double* q = nullptr;
  int64_t q_cap = 0;
  int32_t* cnt = nullptr;
  int64_t cnt_cap = 0;
  int64_t* off = nullptr;
  int64_t off_cap = 0;
  int64_t* idx = nullptr;
  int64_t idx_cap = 0;
  long long* status = nullptr;
};

template <typename T>
int grow(T** p, int64_t* cap, int64_t need) {
  if (*p && *cap >= need) return DP_OK;
  if (*p) DP_HIP_RET(hipFree(*p));
  *p = nullptr;
  *cap = 0;
  const int64_t n = std::max<int64_t>(need, 1);
  DP_HIP_RET(hipMalloc((void**)p, sizeof(T) * (size_t)n));
  *cap = n;
  return DP_OK;
}

}  // namespace

// ---- the scratch of the allocation-free entry point -------------------------
// Per-query hit counts (int32, DP_NOALLOC_MAX_QUERIES) followed by two status
// words (total hits, longest list).  One copy per device, for the life of
// the process; every call on a device reuses it.
#define DP_NOALLOC_COUNTS_BYTES ((size_t)DP_NOALLOC_MAX_QUERIES * sizeof(int32_t))
#define DP_NOALLOC_SCRATCH_BYTES (DP_NOALLOC_COUNTS_BYTES + 256)
static __device__ __align__(256) unsigned char scratch[DP_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` per device
}  // namespace

// Called by dp_create with `device` current: resolves the device address of
// `scratch` once per device.
int dp_noalloc_init_device(int device) {
  if (device < 0 || device >= kMaxDevices) {
    dp_set_error("device index beyond the scratch table");
    return DP_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return DP_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    dp_set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return DP_ERR_HIP;
  }
  g_scratch_ptr[device] = ptr;
  return DP_OK;
}

void dp_radius_free(dp_ctx* c) {
  if (!c->radius) return;
  radius_bufs* b = (radius_bufs*)c->radius;
  (void)hipFree(b->q);
  (void)hipFree(b->cnt);
  (void)hipFree(b->off);
  (void)hipFree(b->idx);
  (void)hipFree(b->status);
  delete b;
  c->radius = nullptr;
}

extern "C" {

int dp_visited_radius(dp_ctx* c, const double* queries, int64_t n, double radius, int64_t* out_offsets,
                      int64_t* out_idx, int64_t idx_cap) {
  if (!c || (n > 0 && !queries) || !out_offsets || n < 0 || idx_cap < 0 || (idx_cap > 0 && !out_idx)) {
    dp_set_error("invalid argument");
    return DP_ERR_ARG;
  }
  if (!(radius >= 0.0) || !std::isfinite(radius)) {
    dp_set_error("radius must be finite and >= 0");
    return DP_ERR_ARG;
  }
  if (!c->have_world) {
    dp_set_error("upload the world before dp_visited_radius (the grid spans its bounds)");
    return DP_ERR_STATE;
  }
  DP_HIP_RET(hipSetDevice(c->device));
  out_offsets[0] = 0;
  if (n == 0) return DP_OK;
  hipStream_t s = c->stream;
  // A grid built for a radius R >= radius has cells wide enough for it.
  if (c->grid_dirty || !(c->grid_radius >= radius)) {
    int rc = dp_launch_visited_rebuild(c, radius, s);
    if (rc) return rc;
  }
  if (c->nv == 0) {
    for (int64_t q = 0; q <= n; ++q) out_offsets[q] = 0;
    return DP_OK;
  }
  if (!c->radius) c->radius = new radius_bufs();
  radius_bufs* B = (radius_bufs*)c->radius;
  const int nx = c->desc.nx;
  int rc;
  DP_HIP_RET(hipStreamSynchronize(s));  // buffers may grow below
  if ((rc = grow(&B->q, &B->q_cap, n * nx))) return rc;
  if ((rc = grow(&B->cnt, &B->cnt_cap, n))) return rc;
  if ((rc = grow(&B->off, &B->off_cap, n + 1))) return rc;
  if (!B->status) DP_HIP_RET(hipMalloc((void**)&B->status, sizeof(long long) * 2));
  DP_HIP_RET(hipMemcpyAsync(B->q, queries, sizeof(double) * n * nx, hipMemcpyHostToDevice, s));
  if ((rc = launch_radius(c, B->q, n, radius, B->cnt, B->off, nullptr, 0, B->status, false, s))) return rc;
  long long st[2];
  DP_HIP_RET(hipMemcpyAsync(st, B->status, sizeof(st), hipMemcpyDeviceToHost, s));
  DP_HIP_RET(hipStreamSynchronize(s));
  const int64_t keep = std::min<int64_t>(st[0], idx_cap);
  if (keep > 0) {
    if ((rc = grow(&B->idx, &B->idx_cap, keep))) return rc;
    if ((rc = launch_radius(c, B->q, n, radius, B->cnt, B->off, B->idx, keep, B->status, true, s))) return rc;
  }
  DP_HIP_RET(hipMemcpyAsync(out_offsets, B->off, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, s));
  if (keep > 0) DP_HIP_RET(hipMemcpyAsync(out_idx, B->idx, sizeof(int64_t) * keep, hipMemcpyDeviceToHost, s));
  DP_HIP_RET(hipStreamSynchronize(s));
  // lists longer than the LDS sort are ordered here
  if (st[1] > DP_RADIUS_SORT_MAX)
    for (int64_t q = 0; q < n; ++q) {
      const int64_t b = out_offsets[q], e = std::min<int64_t>(out_offsets[q + 1], keep);
      if (e - b > DP_RADIUS_SORT_MAX) std::sort(out_idx + b, out_idx + e);
    }
  if (st[0] > idx_cap) {
    dp_set_error("radius query output capacity exceeded");
    return DP_ERR_CAPACITY;
  }
  return DP_OK;
}

int dp_visited_radius_noalloc(dp_ctx* c, const double* d_queries, int64_t n, double radius, int64_t* d_offsets,
                              int64_t* d_idx, int64_t idx_cap, void* stream) {
  if (!c || (n > 0 && !d_queries) || !d_offsets || n < 0 || idx_cap < 0 || (idx_cap > 0 && !d_idx)) {
    dp_set_error("invalid argument");
    return DP_ERR_ARG;
  }
  if (n > DP_NOALLOC_MAX_QUERIES) {
    dp_set_error("dp_visited_radius_noalloc takes at most " + std::to_string(DP_NOALLOC_MAX_QUERIES) +
                 " queries per call");
    return DP_ERR_ARG;
  }
  if (!(radius >= 0.0) || !std::isfinite(radius)) {
    dp_set_error("radius must be finite and >= 0");
    return DP_ERR_ARG;
  }
  if (c->nv > 0 && (c->grid_dirty || !(c->grid_radius >= radius))) {
    dp_set_error("the visited grid is not built for this radius: call dp_visited_radius or dp_expand once "
                 "after the last change of the visited set");
    return DP_ERR_STATE;
  }
  if (c->device < 0 || c->device >= kMaxDevices || !g_scratch_ptr[c->device]) {
    dp_set_error("scratch address of the device not resolved");
    return DP_ERR_STATE;
  }
  DP_HIP_RET(hipSetDevice(c->device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  unsigned char* base = (unsigned char*)g_scratch_ptr[c->device];
  int32_t* counts = (int32_t*)base;
  long long* status = (long long*)(base + DP_NOALLOC_COUNTS_BYTES);
  if (n == 0 || c->nv == 0) {
    // every list is empty: offsets of zero counts
    DP_HIP_RET(hipMemsetAsync(d_offsets, 0, sizeof(int64_t) * (n + 1), s));
    DP_HIP_RET(hipStreamSynchronize(s));
    return DP_OK;
  }
  int rc = launch_radius(c, d_queries, n, radius, counts, d_offsets, nullptr, 0, status, false, s);
  if (rc) return rc;
  rc = launch_radius(c, d_queries, n, radius, counts, d_offsets, d_idx, idx_cap, status, true, s);
  if (rc) return rc;
  DP_HIP_RET(hipMemcpyAsync(c->h_pinned, status, sizeof(long long) * 2, hipMemcpyDeviceToHost, s));
  DP_HIP_RET(hipStreamSynchronize(s));
  const long long total = c->h_pinned[0], longest = c->h_pinned[1];
  if (total > idx_cap) {
    dp_set_error("radius query output capacity exceeded (entries beyond idx_cap not written)");
    return DP_ERR_CAPACITY;
  }
  if (longest > DP_RADIUS_SORT_MAX) {
    dp_set_error("a query has more than 4096 visited states within the radius; its list is not sorted");
    return DP_ERR_CAPACITY;
  }
  return DP_OK;
}

}  // extern "C"
