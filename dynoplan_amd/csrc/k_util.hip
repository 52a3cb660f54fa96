// k_util.hip -- fills and exclusive scans used by the expansion pipeline.
// No hipMemset on the expansion stream: every clear is a kernel, so the
// pipeline stays capturable in a HIP graph.
#include <hip/hip_runtime.h>

#include "dp_internal.h"

namespace {

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v;
}

__global__ void k_zero_counters(long long* c, int n) {
  if (threadIdx.x < n) c[threadIdx.x] = 0;
}

// Block-wide inclusive scan of one value per thread (blockDim = 1024).
__device__ __forceinline__ long long block_inclusive_scan(long long v, long long* s) {
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    long long u = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += u;
    __syncthreads();
  }
  long long r = s[t];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(1024) void k_scan_small(int32_t* data, int64_t n, int32_t* total, int64_t cap,
                                                     long long* counters, int slot) {
  __shared__ long long s[1024];
  const int t = threadIdx.x;
  const int64_t chunk = (n + 1023) / 1024;
  const int64_t b = t * chunk;
  const int64_t e = b + chunk < n ? b + chunk : n;
  long long sum = 0;
  for (int64_t k = b; k < e; ++k) sum += data[k];
  long long run = block_inclusive_scan(sum, s) - sum;
  for (int64_t k = b; k < e; ++k) {
    int32_t v = data[k];
    data[k] = (int32_t)run;
    run += v;
  }
  if (t == 1023) {
    long long tot = run;  // thread 1023 owns the last chunk
    data[n] = (int32_t)tot;
    if (total) *total = (int32_t)(tot < cap ? tot : cap);
    if (slot >= 0) counters[slot] = tot;
    if (tot > cap) counters[DPC_OVERFLOW] = 1;
  }
}

constexpr int kTileThreads = 256;
constexpr int kTileItems = 16;
constexpr int kTile = kTileThreads * kTileItems;  // 4096

// Pass 1: per-tile exclusive scan in place, tile sums to tiles[].
__global__ __launch_bounds__(kTileThreads) void k_scan_tiles(int32_t* data, const int32_t* n_dev, int64_t n_host,
                                                             int32_t* tiles) {
  __shared__ long long s[kTileThreads];
  const int64_t n = n_dev ? (int64_t)*n_dev : n_host;
  const int64_t base = (int64_t)blockIdx.x * kTile;
  if (base >= n) {
    if (threadIdx.x == 0) tiles[blockIdx.x] = 0;
    return;
  }
  const int64_t b = base + (int64_t)threadIdx.x * kTileItems;
  int32_t v[kTileItems];
  long long sum = 0;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    v[k] = b + k < n ? data[b + k] : 0;
    sum += v[k];
  }
  // block scan (blockDim = 256)
  const int t = threadIdx.x;
  s[t] = sum;
  __syncthreads();
  for (int off = 1; off < kTileThreads; off <<= 1) {
    long long u = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += u;
    __syncthreads();
  }
  long long run = s[t] - sum;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    if (b + k < n) data[b + k] = (int32_t)run;
    run += v[k];
  }
  if (t == kTileThreads - 1) tiles[blockIdx.x] = (int32_t)s[t];
}

// Pass 3: add tile offsets; write data[n] = total.
__global__ __launch_bounds__(kTileThreads) void k_scan_add(int32_t* data, const int32_t* n_dev, int64_t n_host,
                                                          const int32_t* tiles, int64_t ntiles, long long* counters,
                                                          int slot) {
  const int64_t n = n_dev ? (int64_t)*n_dev : n_host;
  const int64_t base = (int64_t)blockIdx.x * kTile;
  if (base < n) {
    const int32_t off = tiles[blockIdx.x];
    const int64_t b = base + (int64_t)threadIdx.x * kTileItems;
#pragma unroll
    for (int k = 0; k < kTileItems; ++k)
      if (b + k < n) data[b + k] += off;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int32_t tot = tiles[ntiles];
    data[n] = tot;
    if (slot >= 0) counters[slot] = tot;
  }
}

}  // namespace

void dp_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return;
  k_fill_i32<<<dp_grid_blocks(n, 256, 4096), 256, 0, s>>>(p, n, v);
}

void dp_zero_counters(long long* c, hipStream_t s) { k_zero_counters<<<1, 64, 0, s>>>(c, DPC_COUNT); }

void dp_scan_small(int32_t* data, int64_t n, int32_t* total, int64_t cap, long long* counters, int slot,
                   hipStream_t s) {
  k_scan_small<<<1, 1024, 0, s>>>(data, n, total, cap, counters, slot);
}

int dp_scan_large(dp_ctx* c, int32_t* data, const int32_t* n_dev, int64_t n_max, long long* counters, int slot,
                  hipStream_t s) {
  const int64_t ntiles = (n_max + kTile - 1) / kTile;
  int rc = dp_ensure_i32(&c->d_tiles, &c->tiles_cap, ntiles + 1);
  if (rc) return rc;
  k_scan_tiles<<<(int)ntiles, kTileThreads, 0, s>>>(data, n_dev, n_max, c->d_tiles);
  k_scan_small<<<1, 1024, 0, s>>>(c->d_tiles, ntiles, nullptr, INT64_MAX, counters, -1);
  k_scan_add<<<(int)ntiles, kTileThreads, 0, s>>>(data, n_dev, n_max, c->d_tiles, ntiles, counters, slot);
  DP_HIP_RET(hipGetLastError());
  return DP_OK;
}

// Growth frees a buffer earlier kernels of the batch may still read, so the
// device is drained first (growth happens on the first batches only).
int dp_ensure_i32(int32_t** p, int64_t* cap, int64_t need) {
  if (*cap >= need && *p) return DP_OK;
  dp_bump_epoch();
  DP_HIP_RET(hipDeviceSynchronize());
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  DP_HIP_RET(hipMalloc(p, sizeof(int32_t) * (size_t)(need > 0 ? need : 1)));
  *cap = need;
  return DP_OK;
}

int dp_ensure_f64(double** p, int64_t* cap, int64_t need) {
  if (*cap >= need && *p) return DP_OK;
  dp_bump_epoch();
  DP_HIP_RET(hipDeviceSynchronize());
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  DP_HIP_RET(hipMalloc(p, sizeof(double) * (size_t)(need > 0 ? need : 1)));
  *cap = need;
  return DP_OK;
}

int dp_ensure_u64(uint64_t** p, int64_t* cap, int64_t need) {
  if (*cap >= need && *p) return DP_OK;
  dp_bump_epoch();
  DP_HIP_RET(hipDeviceSynchronize());
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  DP_HIP_RET(hipMalloc(p, sizeof(uint64_t) * (size_t)(need > 0 ? need : 1)));
  *cap = need;
  return DP_OK;
}
