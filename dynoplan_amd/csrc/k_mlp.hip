// k_mlp.hip -- the learned collision-feasibility surrogate.
//
// A conservative pre-filter ahead of the exact check (no counterpart in the
// reference): an MLP scores each (node, primitive) pair from a 32x32
// occupancy patch around the node (0.1 m cells) and a learned 64-d embedding
// of the primitive; pairs scoring below a threshold calibrated on held-out
// exact-checker labels are rejected, every other pair is checked exactly.
//
//   x = [patch(1024) | emb(64)] -> 256 -> 256 -> 256 -> 1   (ReLU, logit)
//
// Inference factors the first layer: A_i = W1p patch_i once per node,
// B_j = W1e emb_j + b1 once per primitive, h1 = relu(A_i + B_j) per pair;
// layers 2-3 run on BF16 MFMA (v_mfma_f32_32x32x8bf16_1k) with FP32
// accumulation, one wave per 32 pairs.  Training uses the generic MFMA GEMM
// below for forward and backward, FP32 master weights and Adam.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "dp_internal.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

// Round-to-nearest-even FP32 -> BF16 (bit pattern), NaN kept quiet.
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---- generic GEMM: C(m,n) = sum_k A(m,k) B(k,n) -------------------------------------
// A, B are BF16 with arbitrary strides; C is FP32.  One wave per 32x32
// output tile, 4 waves per workgroup (64x64).  MFMA 32x32x8 lane layout:
//   A: lane l holds A[l%32][4*(l/32) + i], i = 0..3
//   B: lane l holds B[4*(l/32) + i][l%32]
//   C: acc[j] of lane l is C[8*(j/4) + 4*(l/32) + (j%4)][l%32]
struct gemm_args {
  const uint16_t* A;
  int64_t sam, sak;
  const uint16_t* B;
  int64_t sbk, sbn;
  float* C;
  int64_t ldc;
  uint16_t* Cb;  // optional BF16 copy of the epilogue output (ldc)
  const float* bias;
  int32_t M, N, K;
  int32_t relu;
  int32_t accumulate;  // C += result
};

__global__ __launch_bounds__(256) void k_gemm(gemm_args g) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int m0 = blockIdx.y * 64 + (wave / 2) * 32;
  const int n0 = blockIdx.x * 64 + (wave % 2) * 32;
  if (m0 >= g.M || n0 >= g.N) return;
  const int am = m0 + lane % 32;
  const int bn = n0 + lane % 32;
  const int kq = 4 * (lane / 32);
  f32x16 acc = {};
  for (int k0 = 0; k0 < g.K; k0 += 8) {
    bf16x4 a, b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = k0 + kq + i;
      const bool ka = am < g.M && k < g.K;
      const bool kb = bn < g.N && k < g.K;
      a[i] = ka ? (short)g.A[(int64_t)am * g.sam + (int64_t)k * g.sak] : (short)0;
      b[i] = kb ? (short)g.B[(int64_t)k * g.sbk + (int64_t)bn * g.sbn] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int m = m0 + 8 * (j / 4) + 4 * (lane / 32) + (j % 4);
    const int n = n0 + lane % 32;
    if (m >= g.M || n >= g.N) continue;
    float v = acc[j];
    float* c = g.C + (int64_t)m * g.ldc + n;
    if (g.accumulate) v += *c;
    if (g.bias) v += g.bias[n];
    if (g.relu) v = v > 0.f ? v : 0.f;
    *c = v;
    if (g.Cb) g.Cb[(int64_t)m * g.ldc + n] = f2bf(v);
  }
}

// ---- occupancy patches --------------------------------------------------------------
// patch cell (r, c) of node i covers [x + (c - 16) h, x + (c - 15) h) x
// [y + (r - 16) h, y + (r - 15) h); its value is 1 when the cell centre lies
// in an obstacle box (planar AABBs), else 0.  One workgroup per node.
__global__ __launch_bounds__(256) void k_patch(const double* __restrict__ nodes, int32_t Ncap, int32_t N,
                                               const double* __restrict__ obs, int32_t nobs, double cell,
                                               uint16_t* __restrict__ patch) {
  // Obstacles that can contain a cell centre of this patch go to LDS first
  // (a superset: the exact per-cell test below decides).
  __shared__ double cand[4 * DP_MAX_OBS];
  __shared__ int ncand;
  const int32_t i = blockIdx.x;
  if (i >= N) return;
  const double x = nodes[i], y = nodes[(int64_t)Ncap + i];
  if (threadIdx.x == 0) ncand = 0;
  __syncthreads();
  const double ext = 15.5 * cell + 1e-9;
  for (int o = threadIdx.x; o < nobs; o += blockDim.x) {
    const double* b = obs + 4 * o;
    if (fabs(x - b[0]) <= b[2] + ext && fabs(y - b[1]) <= b[3] + ext) {
      const int k = atomicAdd(&ncand, 1);
      for (int c = 0; c < 4; ++c) cand[4 * k + c] = b[c];
    }
  }
  __syncthreads();
  const int nc = ncand;
  for (int t = threadIdx.x; t < 1024; t += blockDim.x) {
    const int r = t / 32, c = t % 32;
    const double px = x + ((double)c - 15.5) * cell;
    const double py = y + ((double)r - 15.5) * cell;
    bool occ = false;
    for (int o = 0; o < nc && !occ; ++o) {
      const double* b = cand + 4 * o;
      occ = fabs(px - b[0]) <= b[2] && fabs(py - b[1]) <= b[3];
    }
    patch[(int64_t)i * 1024 + t] = occ ? 0x3f80 : 0;  // 1.0 / 0.0 in BF16
  }
}

// ---- fused pair inference -----------------------------------------------------------
struct infer_args {
  const int32_t* pnode;
  const int32_t* pprim;
  const int32_t* total;
  const float* An;   // [N][256] W1p * patch per node
  const float* Bp;   // [P][256] W1e * emb + b1 per primitive
  const uint16_t* W2;  // [256][256] BF16 (out, in)
  const float* b2;
  const uint16_t* W3;
  const float* b3;
  const float* w4;   // [256]
  const float* b4;   // [1]
  float threshold;
  int32_t reject;    // write DP_V_SURROGATE_REJECT below the threshold
  float* score;      // [pair_cap] or null
  int8_t* verdict;
  int32_t* firstk;
  int32_t* goal_out;
  int8_t* res;
  long long* counters;
};

// Inference: 8 waves per workgroup, 32 pairs per wave.  The weights of a
// layer are staged through LDS in K-slices of 32 (256 rows x 64 B, rows
// padded to 72 B against bank conflicts) and shared by the 8 waves, so each
// slice read from L2 serves 256 pairs; activations stay in one 32x256 BF16
// LDS tile per wave.  146 KB of LDS: one workgroup per CU.
constexpr int kInferWaves = 8;
constexpr int kSlice = 32;
constexpr int kSliceStride = kSlice + 4;  // bf16 per staged row

__device__ __forceinline__ void stage_slice(const uint16_t* __restrict__ W, int ks, uint16_t* sw) {
  // row n = t / 2, bytes [32 (t % 2), +32) of W[n][ks .. ks + 32)
  const int t = threadIdx.x;
  const int n = t >> 1, half = t & 1;
  const uint2* src = reinterpret_cast<const uint2*>(W + (int64_t)n * 256 + ks + half * 16);
  uint2* dst = reinterpret_cast<uint2*>(sw + n * kSliceStride + half * 16);
#pragma unroll
  for (int k = 0; k < 4; ++k) dst[k] = src[k];
}

// out[m][n] = relu(sum_k in[m][k] W[n][k] + bias[n]) for this wave's 32 rows
// x 256 columns; `first`: the input is relu(An + Bp) formed on the fly.
__device__ __forceinline__ void mlp_layer(const uint16_t* __restrict__ W, const float* __restrict__ bias,
                                          uint16_t* act, uint16_t* sw, int lane, const float* An_row,
                                          const float* Bp_row, bool first) {
  f32x16 acc[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) acc[t] = f32x16{};
  const int kq = 4 * (lane / 32);
  for (int ks = 0; ks < 256; ks += kSlice) {
    __syncthreads();  // the previous slice has been used by every wave
    stage_slice(W, ks, sw);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kSlice; kk += 8) {
      const int k0 = ks + kk;
      bf16x4 a;
      if (first) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float v = An_row[k0 + kq + i] + Bp_row[k0 + kq + i];
          a[i] = (short)f2bf(v > 0.f ? v : 0.f);
        }
      } else {
        a = *reinterpret_cast<const bf16x4*>(act + (lane % 32) * 256 + k0 + kq);
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(sw + (t * 32 + lane % 32) * kSliceStride + kk + kq);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc[t], 0, 0, 0);
      }
    }
  }
  // every lane has read its inputs (the MFMAs above consumed them): the
  // tile is overwritten in place
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int m = 8 * (j / 4) + 4 * (lane / 32) + (j % 4);
      const int n = t * 32 + lane % 32;
      float v = acc[t][j] + bias[n];
      act[m * 256 + n] = f2bf(v > 0.f ? v : 0.f);
    }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64 * kInferWaves) void k_infer(infer_args a) {
  __shared__ uint16_t act[kInferWaves][32 * 256];
  __shared__ uint16_t sw[256 * kSliceStride];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int32_t tot = *a.total;
  long long n_rej = 0;
  constexpr int kTile = kInferWaves * 32;
  // every wave of the workgroup runs the same number of tiles (the layer
  // loops synchronise the workgroup)
  for (int64_t t0 = (int64_t)blockIdx.x * kTile; t0 < tot; t0 += (int64_t)gridDim.x * kTile) {
    const int64_t q0 = t0 + wave * 32;
    const int64_t qm = q0 + lane % 32;
    const bool vm = qm < tot;
    const int64_t qs = vm ? qm : (tot - 1);  // rows past the end repeat the last pair, unused
    const int32_t i = a.pnode[qs];
    const int32_t j = a.pprim[qs];
    mlp_layer(a.W2, a.b2, act[wave], sw, lane, a.An + (int64_t)i * 256, a.Bp + (int64_t)j * 256, true);
    mlp_layer(a.W3, a.b3, act[wave], sw, lane, nullptr, nullptr, false);
    if (lane < 32 && vm) {
      float z = a.b4[0];
      const uint16_t* h = act[wave] + lane * 256;
      for (int n = 0; n < 256; ++n) z += bf2f(h[n]) * a.w4[n];
      if (a.score) a.score[qm] = z;
      if (a.reject && z < a.threshold) {
        a.verdict[qm] = DP_V_SURROGATE_REJECT;
        a.firstk[qm] = -1;
        a.goal_out[qm] = -1;
        a.res[qm] = DP_R_NONE;
        ++n_rej;
      } else {
        a.verdict[qm] = -1;  // to be checked exactly
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  for (int o = 32; o > 0; o >>= 1) n_rej += __shfl_xor(n_rej, o);
  if (lane == 0 && n_rej) atomicAdd((unsigned long long*)&a.counters[DPC_SURR], (unsigned long long)n_rej);
}

// ---- training helpers ---------------------------------------------------------------
__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    y[t] = f2bf(x[t]);
}

// x_b = [patch(node_b) | emb(prim_b)] for a minibatch of sample ids.
__global__ void k_gather_batch(const int32_t* __restrict__ sample, const int32_t* __restrict__ snode,
                               const int32_t* __restrict__ sprim, const uint16_t* __restrict__ patches,
                               const uint16_t* __restrict__ emb, int32_t B, uint16_t* __restrict__ X) {
  const int32_t b = blockIdx.x;
  if (b >= B) return;
  const int32_t s = sample[b];
  const int32_t i = snode[s], j = sprim[s];
  for (int t = threadIdx.x; t < 1088; t += blockDim.x)
    X[(int64_t)b * 1088 + t] = t < 1024 ? patches[(int64_t)i * 1024 + t] : emb[(int64_t)j * 64 + (t - 1024)];
}

// logits z[b] = h3[b] . w4 + b4; loss = mean BCE; dz = (sigmoid(z) - y) / B.
__global__ void k_head_fwd_bwd(const float* __restrict__ h3, const float* __restrict__ w4, const float* __restrict__ b4,
                               const float* __restrict__ y, int32_t B, float* __restrict__ z, float* __restrict__ dz,
                               float* __restrict__ loss) {
  for (int32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
    float s = b4[0];
    for (int n = 0; n < 256; ++n) s += h3[(int64_t)b * 256 + n] * w4[n];
    z[b] = s;
    const float p = 1.f / (1.f + expf(-s));
    dz[b] = (p - y[b]) / (float)B;
    const float l = fmaxf(s, 0.f) - s * y[b] + log1pf(expf(-fabsf(s)));
    atomicAdd(loss, l / (float)B);
  }
}

// dz3[b][n] = dz[b] w4[n] (h3 > 0) (BF16); dw4[n] += sum_b dz[b] h3[b][n];
// db3[n] += sum_b dz3[b][n]; db4 += sum_b dz[b]
__global__ void k_head_grad(const float* __restrict__ h3, const float* __restrict__ w4, const float* __restrict__ dz,
                            int32_t B, uint16_t* __restrict__ dh3b, float* __restrict__ dw4, float* __restrict__ db3,
                            float* __restrict__ db4) {
  const int n = threadIdx.x;  // 256 threads
  float acc = 0.f, accb = 0.f;
  for (int32_t b = blockIdx.x; b < B; b += gridDim.x) {
    const float h = h3[(int64_t)b * 256 + n];
    acc += dz[b] * h;
    const float g = h > 0.f ? dz[b] * w4[n] : 0.f;
    accb += g;
    dh3b[(int64_t)b * 256 + n] = f2bf(g);
  }
  atomicAdd(&dw4[n], acc);
  atomicAdd(&db3[n], accb);
  if (n == 0) {
    float s = 0.f;
    for (int32_t b = blockIdx.x; b < B; b += gridDim.x) s += dz[b];
    atomicAdd(db4, s);
  }
}

// dh (FP32, from GEMM dY W) masked by relu'(h) -> BF16; bias grad += column sums.
__global__ void k_relu_back(const float* __restrict__ dh, const float* __restrict__ h, int32_t B, int32_t F,
                            uint16_t* __restrict__ dhb, float* __restrict__ dbias) {
  const int n = threadIdx.x + blockIdx.y * blockDim.x;
  if (n >= F) return;
  float acc = 0.f;
  for (int32_t b = blockIdx.x; b < B; b += gridDim.x) {
    const float g = h[(int64_t)b * F + n] > 0.f ? dh[(int64_t)b * F + n] : 0.f;
    dhb[(int64_t)b * F + n] = f2bf(g);
    acc += g;
  }
  if (dbias) atomicAdd(&dbias[n], acc);
}

// gradient of the 64 embedding inputs of each sample -> embedding rows
__global__ void k_emb_scatter(const float* __restrict__ dX, const int32_t* __restrict__ sample,
                              const int32_t* __restrict__ sprim, int32_t B, float* __restrict__ demb) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)B * 64;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t b = (int32_t)(t / 64), c = (int32_t)(t % 64);
    atomicAdd(&demb[(int64_t)sprim[sample[b]] * 64 + c], dX[(int64_t)b * 64 + c]);
  }
}

__global__ void k_adam(float* __restrict__ w, float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.f - b1) * gt;
    const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    w[t] -= lr * (mt / bc1) / (sqrtf(vt / bc2) + eps);
    g[t] = 0.f;
  }
}

__global__ void k_gather_labels(const int32_t* __restrict__ sid, const float* __restrict__ label, int32_t B,
                                float* __restrict__ y) {
  for (int32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) y[b] = label[sid[b]];
}

__global__ void k_fill_f32(float* p, int64_t n, float v) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) p[t] = v;
}

// sample ids of a minibatch: a counter-based hash of (step, b)
__global__ void k_sample_ids(uint64_t seed, int64_t step, int32_t B, int32_t nsamples, int32_t* __restrict__ out) {
  for (int32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
    uint64_t z = seed ^ ((uint64_t)step * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)b * 0xC2B2AE3D27D4EB4Full);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    out[b] = (int32_t)(z % (uint64_t)nsamples);
  }
}

}  // namespace

// ==========================================================================
// host side
// ==========================================================================

// Parameter block offsets (FP32 master vector).
struct mlp_layout {
  int64_t W1, b1, W2, b2, W3, b3, w4, b4, total;
};
static mlp_layout layout_of() {
  mlp_layout L;
  int64_t o = 0;
  L.W1 = o; o += 256 * 1088;
  L.b1 = o; o += 256;
  L.W2 = o; o += 256 * 256;
  L.b2 = o; o += 256;
  L.W3 = o; o += 256 * 256;
  L.b3 = o; o += 256;
  L.w4 = o; o += 256;
  L.b4 = o; o += 1;
  L.total = o;
  return L;
}

int64_t dp_mlp_param_count() { return layout_of().total; }

static void gemm(hipStream_t s, const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
                 int64_t sbn, float* C, int64_t ldc, uint16_t* Cb, const float* bias, int M, int N, int K, bool relu,
                 bool accumulate) {
  gemm_args g{A, sam, sak, B, sbk, sbn, C, ldc, Cb, bias, M, N, K, relu ? 1 : 0, accumulate ? 1 : 0};
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  k_gemm<<<grid, 256, 0, s>>>(g);
}

static int ensure_u16(uint16_t** p, int64_t* cap, int64_t need) {
  if (*cap >= need && *p) return DP_OK;
  dp_bump_epoch();
  DP_HIP_RET(hipDeviceSynchronize());
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  DP_HIP_RET(hipMalloc(p, sizeof(uint16_t) * (size_t)need));
  *cap = need;
  return DP_OK;
}

static int ensure_f32(float** p, int64_t* cap, int64_t need) {
  if (*cap >= need && *p) return DP_OK;
  dp_bump_epoch();
  DP_HIP_RET(hipDeviceSynchronize());
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  DP_HIP_RET(hipMalloc(p, sizeof(float) * (size_t)need));
  *cap = need;
  return DP_OK;
}

int dp_mlp_alloc(dp_ctx* c, int32_t P, uint64_t seed) {
  dp_mlp& m = c->mlp;
  const mlp_layout L = layout_of();
  int rc;
  if ((rc = ensure_f32(&m.w, &m.w_cap, L.total))) return rc;
  if ((rc = ensure_f32(&m.g, &m.g_cap, L.total))) return rc;
  if ((rc = ensure_f32(&m.m1, &m.m1_cap, L.total))) return rc;
  if ((rc = ensure_f32(&m.m2, &m.m2_cap, L.total))) return rc;
  if ((rc = ensure_u16(&m.wb, &m.wb_cap, L.total))) return rc;
  if ((rc = ensure_f32(&m.emb, &m.emb_cap, (int64_t)P * 64 * 4))) return rc;  // emb, grad, m, v
  if ((rc = ensure_u16(&m.embb, &m.embb_cap, (int64_t)P * 64))) return rc;
  // He-style uniform init from a seeded stream (host), biases zero.
  std::vector<float> w((size_t)L.total, 0.f), e((size_t)P * 64 * 4, 0.f);
  uint64_t st = seed ? seed : 0x1234567ull;
  auto unif = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (float)((double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0);
  };
  auto fill = [&](int64_t off, int64_t n, float scale) {
    for (int64_t k = 0; k < n; ++k) w[off + k] = scale * unif();
  };
  fill(L.W1, 256 * 1088, std::sqrt(6.f / 1088.f));
  fill(L.W2, 256 * 256, std::sqrt(6.f / 256.f));
  fill(L.W3, 256 * 256, std::sqrt(6.f / 256.f));
  fill(L.w4, 256, std::sqrt(6.f / 256.f));
  for (int64_t k = 0; k < (int64_t)P * 64; ++k) e[k] = 0.1f * unif();
  DP_HIP_RET(hipMemcpy(m.w, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
  DP_HIP_RET(hipMemcpy(m.emb, e.data(), sizeof(float) * e.size(), hipMemcpyHostToDevice));
  DP_HIP_RET(hipMemset(m.g, 0, sizeof(float) * L.total));
  DP_HIP_RET(hipMemset(m.m1, 0, sizeof(float) * L.total));
  DP_HIP_RET(hipMemset(m.m2, 0, sizeof(float) * L.total));
  m.P = P;
  m.step = 0;
  m.ready = false;
  m.cache_valid = false;
  return DP_OK;
}

static int refresh_bf16(dp_ctx* c, hipStream_t s) {
  dp_mlp& m = c->mlp;
  const mlp_layout L = layout_of();
  k_to_bf16<<<dp_grid_blocks(L.total, 256, 2048), 256, 0, s>>>(m.w, m.wb, L.total);
  k_to_bf16<<<dp_grid_blocks((int64_t)m.P * 64, 256, 2048), 256, 0, s>>>(m.emb, m.embb, (int64_t)m.P * 64);
  m.cache_valid = false;
  dp_bump_epoch();
  DP_HIP_RET(hipGetLastError());
  return DP_OK;
}

// Forward + backward of one minibatch X [B][1088] (BF16), labels y [B];
// gradients accumulate into m.g and the embedding gradient block.
int dp_mlp_fwd_bwd(dp_ctx* c, const uint16_t* X, const float* y, int32_t B, const int32_t* sample,
                   const int32_t* sprim, float* loss, hipStream_t s) {
  dp_mlp& m = c->mlp;
  const mlp_layout L = layout_of();
  int rc;
  const int64_t need = (int64_t)B * 256;
  if ((rc = ensure_f32(&m.h1, &m.h1_cap, need))) return rc;
  if ((rc = ensure_f32(&m.h2, &m.h2_cap, need))) return rc;
  if ((rc = ensure_f32(&m.h3, &m.h3_cap, need))) return rc;
  if ((rc = ensure_u16(&m.h1b, &m.h1b_cap, need))) return rc;
  if ((rc = ensure_u16(&m.h2b, &m.h2b_cap, need))) return rc;
  if ((rc = ensure_u16(&m.h3b, &m.h3b_cap, need))) return rc;
  if ((rc = ensure_f32(&m.dh, &m.dh_cap, (int64_t)B * 1088))) return rc;
  if ((rc = ensure_u16(&m.dhb, &m.dhb_cap, need))) return rc;
  if ((rc = ensure_f32(&m.zb, &m.zb_cap, 2 * (int64_t)B))) return rc;
  const uint16_t* Wb = m.wb;
  float* G = m.g;
  // forward: h = relu(X W^T + b)   (W stored [out][in])
  gemm(s, X, 1088, 1, Wb + L.W1, 1, 1088, m.h1, 256, m.h1b, m.w + L.b1, B, 256, 1088, true, false);
  gemm(s, m.h1b, 256, 1, Wb + L.W2, 1, 256, m.h2, 256, m.h2b, m.w + L.b2, B, 256, 256, true, false);
  gemm(s, m.h2b, 256, 1, Wb + L.W3, 1, 256, m.h3, 256, m.h3b, m.w + L.b3, B, 256, 256, true, false);
  float* z = m.zb;
  float* dz = m.zb + B;
  k_head_fwd_bwd<<<dp_grid_blocks(B, 256, 1024), 256, 0, s>>>(m.h3, m.w + L.w4, m.w + L.b4, y, B, z, dz, loss);
  // head: dh3 (BF16, relu-masked), dw4, db4
  k_head_grad<<<dp_grid_blocks(B, 1, 512), 256, 0, s>>>(m.h3, m.w + L.w4, dz, B, m.dhb, G + L.w4, G + L.b3,
                                                        G + L.b4);
  // layer 3: dW3 += dz3^T h2 ; dh2 = dz3 W3 ; dz2 = dh2 (h2 > 0), db2
  gemm(s, m.dhb, 1, 256, m.h2b, 256, 1, G + L.W3, 256, nullptr, nullptr, 256, 256, B, false, true);
  gemm(s, m.dhb, 256, 1, Wb + L.W3, 256, 1, m.dh, 256, nullptr, nullptr, B, 256, 256, false, false);
  k_relu_back<<<dim3(dp_grid_blocks(B, 1, 256), 1), 256, 0, s>>>(m.dh, m.h2, B, 256, m.dhb, G + L.b2);
  // layer 2: dW2 += dz2^T h1 ; dh1 = dz2 W2 ; dz1 = dh1 (h1 > 0), db1
  gemm(s, m.dhb, 1, 256, m.h1b, 256, 1, G + L.W2, 256, nullptr, nullptr, 256, 256, B, false, true);
  gemm(s, m.dhb, 256, 1, Wb + L.W2, 256, 1, m.dh, 256, nullptr, nullptr, B, 256, 256, false, false);
  k_relu_back<<<dim3(dp_grid_blocks(B, 1, 256), 1), 256, 0, s>>>(m.dh, m.h1, B, 256, m.dhb, G + L.b1);
  // layer 1: dW1 += dz1^T X ; embedding inputs: dX[:, 1024:] = dz1 W1[:, 1024:]
  gemm(s, m.dhb, 1, 256, X, 1088, 1, G + L.W1, 1088, nullptr, nullptr, 256, 1088, B, false, true);
  if (sample && sprim) {
    gemm(s, m.dhb, 256, 1, Wb + L.W1 + 1024, 1088, 1, m.dh, 64, nullptr, nullptr, B, 64, 256, false, false);
    k_emb_scatter<<<dp_grid_blocks((int64_t)B * 64, 256, 2048), 256, 0, s>>>(m.dh, sample, sprim, B,
                                                                             m.emb + (int64_t)m.P * 64);
  }
  DP_HIP_RET(hipGetLastError());
  return DP_OK;
}

void dp_mlp_free(dp_ctx* c) {
  dp_mlp& m = c->mlp;
  float* fs[] = {m.w, m.g, m.m1, m.m2, m.emb, m.h1, m.h2, m.h3, m.dh, m.zb, m.An, m.Bp, m.score, m.slabel, m.loss};
  for (float* p : fs)
    if (p) (void)hipFree(p);
  uint16_t* hs[] = {m.wb, m.embb, m.h1b, m.h2b, m.h3b, m.dhb, m.xb, m.patch};
  for (uint16_t* p : hs)
    if (p) (void)hipFree(p);
  int32_t* is[] = {m.snode, m.sprim, m.sid};
  for (int32_t* p : is)
    if (p) (void)hipFree(p);
  m = dp_mlp{};
}

static int patches_and_node_terms(dp_ctx* c, hipStream_t s) {
  dp_mlp& m = c->mlp;
  const mlp_layout L = layout_of();
  int rc;
  if ((rc = ensure_u16(&m.patch, &m.patch_cap, (int64_t)c->N_cap * 1024))) return rc;
  if ((rc = ensure_f32(&m.An, &m.An_cap, (int64_t)c->N_cap * 256))) return rc;
  k_patch<<<c->N, 256, 0, s>>>(c->d_nodes, c->N_cap, c->N, c->d_obs, c->nb, 0.1, m.patch);
  // A_i = W1[:, :1024] patch_i
  gemm(s, m.patch, 1024, 1, m.wb + L.W1, 1, 1088, m.An, 256, nullptr, nullptr, c->N, 256, 1024, false, false);
  DP_HIP_RET(hipGetLastError());
  return DP_OK;
}

int dp_launch_surrogate(dp_ctx* c, bool reject, hipStream_t s) {
  dp_mlp& m = c->mlp;
  const mlp_layout L = layout_of();
  int rc;
  if ((rc = patches_and_node_terms(c, s))) return rc;
  if (!m.cache_valid) {
    if ((rc = ensure_f32(&m.Bp, &m.Bp_cap, (int64_t)m.P * 256))) return rc;
    // B_j = W1[:, 1024:] emb_j + b1
    gemm(s, m.embb, 64, 1, m.wb + L.W1 + 1024, 1, 1088, m.Bp, 256, nullptr, m.w + L.b1, m.P, 256, 64, false, false);
    m.cache_valid = true;
  }
  if ((rc = ensure_f32(&m.score, &m.score_cap, c->pair_cap))) return rc;
  infer_args a{};
  a.pnode = c->d_pnode;
  a.pprim = c->d_pprim;
  a.total = c->d_total;
  a.An = m.An;
  a.Bp = m.Bp;
  a.W2 = m.wb + L.W2;
  a.b2 = m.w + L.b2;
  a.W3 = m.wb + L.W3;
  a.b3 = m.w + L.b3;
  a.w4 = m.w + L.w4;
  a.b4 = m.w + L.b4;
  a.threshold = m.threshold;
  a.reject = reject ? 1 : 0;
  a.score = m.score;
  a.verdict = c->d_verdict;
  a.firstk = c->d_firstk;
  a.goal_out = c->d_goal;
  a.res = c->d_res;
  a.counters = c->d_counters;
  // one workgroup per CU (LDS-bound); tiles of 256 pairs, grid-stride
  const int64_t tiles = ((int64_t)c->N * c->P + 255) / 256;
  k_infer<<<dp_grid_blocks(tiles, 1, 256), 64 * kInferWaves, 0, s>>>(a);
  DP_HIP_RET(hipGetLastError());
  return DP_OK;
}

static int adam_step(dp_ctx* c, float lr, hipStream_t s) {
  dp_mlp& m = c->mlp;
  const mlp_layout L = layout_of();
  m.step += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float bc1 = 1.f - std::pow(b1, (float)m.step), bc2 = 1.f - std::pow(b2, (float)m.step);
  k_adam<<<dp_grid_blocks(L.total, 256, 2048), 256, 0, s>>>(m.w, m.g, m.m1, m.m2, L.total, lr, b1, b2, eps, bc1, bc2);
  const int64_t E = (int64_t)m.P * 64;
  k_adam<<<dp_grid_blocks(E, 256, 2048), 256, 0, s>>>(m.emb, m.emb + E, m.emb + 2 * E, m.emb + 3 * E, E, lr, b1, b2,
                                                        eps, bc1, bc2);
  return refresh_bf16(c, s);
}

extern "C" {

int dp_surrogate_init(dp_ctx* c, uint64_t seed) {
  if (!c || c->P <= 0) {
    dp_set_error("upload primitives before dp_surrogate_init");
    return DP_ERR_STATE;
  }
  if (c->desc.pos_dim != 2) {
    dp_set_error("the occupancy-patch surrogate is defined for planar systems");
    return DP_ERR_ARG;
  }
  DP_HIP_RET(hipSetDevice(c->device));
  int rc = dp_mlp_alloc(c, c->P, seed);
  if (rc) return rc;
  return refresh_bf16(c, c->stream);
}

// Label the pairs of `nodes` with the exact checker (FREE = 1) and train for
// `epochs` passes over them with minibatches of `batch`.  loss_out (optional,
// `epochs` floats) receives the mean loss of each epoch.
int dp_surrogate_train(dp_ctx* c, const double* nodes, int32_t N, const dp_expand_params* params, int32_t epochs,
                       int32_t batch, float lr, uint64_t seed, float* loss_out) {
  if (!c || !nodes || !params || N <= 0 || epochs <= 0 || batch <= 0) {
    dp_set_error("invalid argument");
    return DP_ERR_ARG;
  }
  dp_mlp& m = c->mlp;
  if (!m.w || m.P != c->P) {
    dp_set_error("call dp_surrogate_init first");
    return DP_ERR_STATE;
  }
  dp_expand_params p = *params;
  p.use_surrogate = 0;
  int rc = dp_expand_batch(c, nodes, nullptr, N, &p, nullptr);
  if (rc) return rc;
  hipStream_t s = c->stream;
  int64_t npairs = 0;
  {
    long long h = 0;
    DP_HIP_RET(hipMemcpy(&h, c->d_counters + DPC_PAIRS, sizeof(long long), hipMemcpyDeviceToHost));
    npairs = h;
  }
  if (npairs <= 0) {
    dp_set_error("no applicable pairs in the training nodes");
    return DP_ERR_STATE;
  }
  if ((rc = dp_ensure_i32(&m.snode, &m.s_cap, npairs))) return rc;
  {
    int64_t cap2 = 0;
    if (m.sprim) (void)hipFree(m.sprim);
    m.sprim = nullptr;
    if ((rc = dp_ensure_i32(&m.sprim, &cap2, npairs))) return rc;
    if (m.slabel) (void)hipFree(m.slabel);
    m.slabel = nullptr;
    int64_t cap3 = 0;
    if ((rc = ensure_f32(&m.slabel, &cap3, npairs))) return rc;
  }
  // on the context stream: device-to-device copies do not block the host,
  // and the null stream does not order with the context stream
  DP_HIP_RET(hipMemcpyAsync(m.snode, c->d_pnode, sizeof(int32_t) * npairs, hipMemcpyDeviceToDevice, s));
  DP_HIP_RET(hipMemcpyAsync(m.sprim, c->d_pprim, sizeof(int32_t) * npairs, hipMemcpyDeviceToDevice, s));
  {
    std::vector<int8_t> v((size_t)npairs);
    std::vector<float> y((size_t)npairs);
    DP_HIP_RET(hipMemcpy(v.data(), c->d_verdict, npairs, hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < npairs; ++q) y[q] = v[q] == DP_V_FREE ? 1.f : 0.f;
    DP_HIP_RET(hipMemcpy(m.slabel, y.data(), sizeof(float) * npairs, hipMemcpyHostToDevice));
  }
  // occupancy patches of the training nodes
  if ((rc = ensure_u16(&m.patch, &m.patch_cap, (int64_t)c->N_cap * 1024))) return rc;
  k_patch<<<c->N, 256, 0, s>>>(c->d_nodes, c->N_cap, c->N, c->d_obs, c->nb, 0.1, m.patch);
  if ((rc = ensure_u16(&m.xb, &m.xb_cap, (int64_t)batch * 1088))) return rc;
  if ((rc = dp_ensure_i32(&m.sid, &m.sid_cap, batch))) return rc;
  if (!m.loss) {
    dp_bump_epoch();
    DP_HIP_RET(hipMalloc(&m.loss, sizeof(float) * 2));
  }
  const int64_t steps_per_epoch = (npairs + batch - 1) / batch;
  for (int32_t e = 0; e < epochs; ++e) {
    DP_HIP_RET(hipMemsetAsync(m.loss + 1, 0, sizeof(float), s));
    for (int64_t t = 0; t < steps_per_epoch; ++t) {
      k_sample_ids<<<dp_grid_blocks(batch, 256, 256), 256, 0, s>>>(seed, m.step, batch, (int32_t)npairs, m.sid);
      k_gather_batch<<<batch, 256, 0, s>>>(m.sid, m.snode, m.sprim, m.patch, m.embb, batch, m.xb);
      // labels of the minibatch, gathered into the head's scratch (zb)
      if ((rc = ensure_f32(&m.zb, &m.zb_cap, 3 * (int64_t)batch))) return rc;
      k_gather_labels<<<dp_grid_blocks(batch, 256, 256), 256, 0, s>>>(m.sid, m.slabel, batch, m.zb + 2 * batch);
      rc = dp_mlp_fwd_bwd(c, m.xb, m.zb + 2 * batch, batch, m.sid, m.sprim, m.loss + 1, s);
      if (rc) return rc;
      if ((rc = adam_step(c, lr, s))) return rc;
    }
    if (loss_out) {
      float h[2];
      DP_HIP_RET(hipStreamSynchronize(s));  // the context stream does not order with the null stream
      DP_HIP_RET(hipMemcpy(h, m.loss, sizeof(float) * 2, hipMemcpyDeviceToHost));
      loss_out[e] = h[1] / (float)steps_per_epoch;
    }
  }
  DP_HIP_RET(hipStreamSynchronize(s));
  m.cache_valid = false;
  return DP_OK;
}

}  // extern "C"

extern "C" {

// Score the pairs of `nodes` without rejecting, compare with the exact
// verdicts, and set the threshold to (lowest score of a FREE pair) - margin:
// on these held-out nodes every feasible edge is kept (recall 1).
int dp_surrogate_calibrate(dp_ctx* c, const double* nodes, int32_t N, const dp_expand_params* params, float margin,
                           float tail,
                           dp_surrogate_stats* out) {
  if (!c || !nodes || !params || N <= 0) {
    dp_set_error("invalid argument");
    return DP_ERR_ARG;
  }
  dp_mlp& m = c->mlp;
  if (!m.w) {
    dp_set_error("call dp_surrogate_init / dp_surrogate_train first");
    return DP_ERR_STATE;
  }
  dp_expand_params p = *params;
  p.use_surrogate = 2;  // score only
  int rc = dp_expand_batch(c, nodes, nullptr, N, &p, nullptr);
  if (rc) return rc;
  long long np = 0;
  DP_HIP_RET(hipMemcpy(&np, c->d_counters + DPC_PAIRS, sizeof(long long), hipMemcpyDeviceToHost));
  std::vector<float> sc((size_t)np);
  std::vector<int8_t> v((size_t)np);
  if (np > 0) {
    DP_HIP_RET(hipMemcpy(sc.data(), m.score, sizeof(float) * np, hipMemcpyDeviceToHost));
    DP_HIP_RET(hipMemcpy(v.data(), c->d_verdict, np, hipMemcpyDeviceToHost));
  }
  // the ten lowest scores of feasible pairs
  std::vector<float> low;
  int64_t n_free = 0;
  for (long long q = 0; q < np; ++q)
    if (v[q] == DP_V_FREE) {
      ++n_free;
      low.push_back(sc[q]);
      std::push_heap(low.begin(), low.end());
      if (low.size() > 10) {
        std::pop_heap(low.begin(), low.end());
        low.pop_back();
      }
    }
  std::sort(low.begin(), low.end());
  float tau = -INFINITY;
  if (n_free) {
    const float spread = low.back() - low.front();
    tau = low.front() - std::max(margin, tail * spread);
  }
  int64_t rej = 0, rej_col = 0, n_col = 0;
  for (long long q = 0; q < np; ++q) {
    if (v[q] == DP_V_COLLIDE) ++n_col;
    if (v[q] != DP_V_FREE && sc[q] < tau) {
      ++rej;
      rej_col += v[q] == DP_V_COLLIDE;
    }
  }
  m.threshold = tau;
  m.ready = true;
  dp_bump_epoch();
  if (out) {
    out->threshold = tau;
    out->pairs = np;
    out->feasible = n_free;
    out->infeasible = np - n_free;
    out->rejectable = rej;
    out->colliding = n_col;
    out->rejectable_colliding = rej_col;
  }
  return DP_OK;
}

int dp_surrogate_set_threshold(dp_ctx* c, float tau) {
  if (!c || !c->mlp.w) {
    dp_set_error("no surrogate");
    return DP_ERR_STATE;
  }
  c->mlp.threshold = tau;
  c->mlp.ready = true;
  dp_bump_epoch();
  return DP_OK;
}

// Scores of the last expansion run with the surrogate (n <= pairs).
int dp_surrogate_scores(dp_ctx* c, float* out, int64_t n) {
  if (!c || !out || n < 0 || !c->mlp.score || n > c->mlp.score_cap) {
    dp_set_error("invalid argument or no scores");
    return DP_ERR_ARG;
  }
  DP_HIP_RET(hipDeviceSynchronize());
  DP_HIP_RET(hipMemcpy(out, c->mlp.score, sizeof(float) * n, hipMemcpyDeviceToHost));
  return DP_OK;
}

int64_t dp_surrogate_param_count(void) { return dp_mlp_param_count(); }

// One forward + backward pass of the MLP on explicit inputs (BF16-rounded on
// the device) with the given FP32 parameters; returns the mean BCE loss and
// the parameter gradient (the layout of dp_surrogate_param_count()).
int dp_surrogate_grad(dp_ctx* c, const float* params, const float* X, const float* y, int32_t B, float* loss,
                      float* grad) {
  if (!c || !params || !X || !y || B <= 0 || !loss || !grad) {
    dp_set_error("invalid argument");
    return DP_ERR_ARG;
  }
  dp_mlp& m = c->mlp;
  if (!m.w) {
    dp_set_error("call dp_surrogate_init first");
    return DP_ERR_STATE;
  }
  const mlp_layout L = layout_of();
  hipStream_t s = c->stream;
  DP_HIP_RET(hipSetDevice(c->device));
  DP_HIP_RET(hipDeviceSynchronize());
  DP_HIP_RET(hipMemcpy(m.w, params, sizeof(float) * L.total, hipMemcpyHostToDevice));
  DP_HIP_RET(hipMemset(m.g, 0, sizeof(float) * L.total));
  int rc = refresh_bf16(c, s);
  if (rc) return rc;
  float *xf = nullptr, *yd = nullptr, *ld = nullptr;
  uint16_t* xb = nullptr;
  DP_HIP_RET(hipMalloc(&xf, sizeof(float) * (size_t)B * 1088));
  DP_HIP_RET(hipMalloc(&yd, sizeof(float) * B));
  DP_HIP_RET(hipMalloc(&ld, sizeof(float)));
  DP_HIP_RET(hipMalloc(&xb, sizeof(uint16_t) * (size_t)B * 1088));
  DP_HIP_RET(hipMemcpy(xf, X, sizeof(float) * (size_t)B * 1088, hipMemcpyHostToDevice));
  DP_HIP_RET(hipMemcpy(yd, y, sizeof(float) * B, hipMemcpyHostToDevice));
  DP_HIP_RET(hipMemset(ld, 0, sizeof(float)));
  k_to_bf16<<<dp_grid_blocks((int64_t)B * 1088, 256, 2048), 256, 0, s>>>(xf, xb, (int64_t)B * 1088);
  rc = dp_mlp_fwd_bwd(c, xb, yd, B, nullptr, nullptr, ld, s);
  if (!rc) {
    DP_HIP_RET(hipStreamSynchronize(s));
    DP_HIP_RET(hipMemcpy(loss, ld, sizeof(float), hipMemcpyDeviceToHost));
    DP_HIP_RET(hipMemcpy(grad, m.g, sizeof(float) * L.total, hipMemcpyDeviceToHost));
  }
  (void)hipFree(xf);
  (void)hipFree(yd);
  (void)hipFree(ld);
  (void)hipFree(xb);
  DP_HIP_RET(hipMemset(m.g, 0, sizeof(float) * L.total));
  return rc;
}

int dp_surrogate_get(dp_ctx* c, float* params, float* emb) {
  if (!c || !c->mlp.w) {
    dp_set_error("no surrogate");
    return DP_ERR_STATE;
  }
  DP_HIP_RET(hipSetDevice(c->device));
  DP_HIP_RET(hipDeviceSynchronize());
  if (params) DP_HIP_RET(hipMemcpy(params, c->mlp.w, sizeof(float) * layout_of().total, hipMemcpyDeviceToHost));
  if (emb) DP_HIP_RET(hipMemcpy(emb, c->mlp.emb, sizeof(float) * (size_t)c->mlp.P * 64, hipMemcpyDeviceToHost));
  return DP_OK;
}

// C = A B with A [M][K], B [K][N] FP32 inputs rounded to BF16 (numerics test
// of the MFMA GEMM used by training).
int dp_debug_gemm_bf16(dp_ctx* c, const float* A, const float* B, int32_t M, int32_t N, int32_t K, float* C) {
  if (!c || !A || !B || !C || M <= 0 || N <= 0 || K <= 0) {
    dp_set_error("invalid argument");
    return DP_ERR_ARG;
  }
  DP_HIP_RET(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  float *af, *bf, *cf;
  uint16_t *ab, *bb;
  DP_HIP_RET(hipMalloc(&af, sizeof(float) * (size_t)M * K));
  DP_HIP_RET(hipMalloc(&bf, sizeof(float) * (size_t)K * N));
  DP_HIP_RET(hipMalloc(&cf, sizeof(float) * (size_t)M * N));
  DP_HIP_RET(hipMalloc(&ab, sizeof(uint16_t) * (size_t)M * K));
  DP_HIP_RET(hipMalloc(&bb, sizeof(uint16_t) * (size_t)K * N));
  DP_HIP_RET(hipMemcpy(af, A, sizeof(float) * (size_t)M * K, hipMemcpyHostToDevice));
  DP_HIP_RET(hipMemcpy(bf, B, sizeof(float) * (size_t)K * N, hipMemcpyHostToDevice));
  k_to_bf16<<<dp_grid_blocks((int64_t)M * K, 256, 2048), 256, 0, s>>>(af, ab, (int64_t)M * K);
  k_to_bf16<<<dp_grid_blocks((int64_t)K * N, 256, 2048), 256, 0, s>>>(bf, bb, (int64_t)K * N);
  gemm(s, ab, K, 1, bb, N, 1, cf, N, nullptr, nullptr, M, N, K, false, false);
  DP_HIP_RET(hipStreamSynchronize(s));
  DP_HIP_RET(hipMemcpy(C, cf, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToHost));
  (void)hipFree(af);
  (void)hipFree(bf);
  (void)hipFree(cf);
  (void)hipFree(ab);
  (void)hipFree(bb);
  return DP_OK;
}

}  // extern "C"
