// mlp.hip — next-treelet predictor: BF16 MFMA MLP with FP32 accumulation.
//
// No counterpart in the reference, whose RT unit only has popularity
// heuristics (src/gpgpu-sim/shader.cc:3419-3559).  The model predicts, at a
// treelet transition of a ray, which treelet the ray enters next:
//   x  = [E[cur] (32) | one-hot octant of the ray direction (8) | depth/32 (1) |
//         E[h1] E[h2] E[h3] E[h4] (4 x 32)]           169 inputs, padded to 192
//   h1 = relu(x W1 + b1)       128
//   h2 = relu(h1 W2 + b2)      128
//   q  = h2 W3 + b3            32
//   logit[c] = q . O[c] + o[c] over the output classes (softmax, top-4)
// Every matrix product runs on v_mfma_f32_16x16x32_bf16: operands are rounded
// to BF16 (round to nearest even) as they are loaded, accumulation is FP32.
// Parameters are FP32 master copies updated with Adam.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "rtx.h"

namespace {

constexpr int kEmb = 32;
constexpr int kHid = 128;
constexpr int kIn = 169;
constexpr int kInPad = 192;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---- GEMM: C[M x N] (+)= op(A)[M x K] . op(B)[K x N], FP32 in memory --------
// 256 threads = 4 waves, each wave a 32x32 tile as 2x2 MFMA 16x16 tiles; the
// block tile is 64x64 (2x2 waves) or, for N <= 32, 128x32 (4x1 waves).
// MFMA 16x16x32 operand layout: lane l supplies row/column l % 16 and the 8
// consecutive k values starting at 8 * (l / 16); the result holds rows
// 4 * (l / 16) + i (i < 4) of column l % 16.  With gridDim.z > 1 the K range
// is split across blocks and the partial products are added into C atomically
// (C already holds the addend: zero or the value to accumulate into).
template <bool TA, bool TB, int BN>
__global__ __launch_bounds__(256) void gemm_bf16(int M, int N, int K, int k_chunk, const float* __restrict__ A,
                                                 int lda, const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                 int ldc, int accumulate) {
  constexpr int WAVES_N = BN / 32;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int row0 = blockIdx.y * (128 / WAVES_N) + (wave / WAVES_N) * 32;
  const int col0 = blockIdx.x * BN + (wave % WAVES_N) * 32;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  const int kb = blockIdx.z * k_chunk;
  const int ke = min(K, kb + k_chunk);
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += 32) {
    bf16x8 a[2], b[2];
    for (int i = 0; i < 2; ++i) {
      const int m = row0 + i * 16 + r16;
      for (int t = 0; t < 8; ++t) {
        const int k = k0 + kq + t;
        float v = 0.f;
        if (m < M && k < ke) v = TA ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k];
        a[i][t] = (__bf16)v;
      }
    }
    for (int j = 0; j < 2; ++j) {
      const int n = col0 + j * 16 + r16;
      for (int t = 0; t < 8; ++t) {
        const int k = k0 + kq + t;
        float v = 0.f;
        if (n < N && k < ke) v = TB ? B[(size_t)n * ldb + k] : B[(size_t)k * ldb + n];
        b[j][t] = (__bf16)v;
      }
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
  const bool split = gridDim.z > 1;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const int n = col0 + j * 16 + r16;
      for (int t = 0; t < 4; ++t) {
        const int m = row0 + i * 16 + (lane >> 4) * 4 + t;
        if (m < M && n < N) {
          float* c = C + (size_t)m * ldc + n;
          if (split)
            atomicAdd(c, acc[i][j][t]);
          else
            *c = accumulate ? *c + acc[i][j][t] : acc[i][j][t];
        }
      }
    }
}

__global__ void zero_rows(float* C, int M, int N, int ldc) {
  const size_t total = (size_t)M * N;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < total; k += (size_t)gridDim.x * blockDim.x)
    C[(k / N) * ldc + k % N] = 0.f;
}

// ---- elementwise and reductions ---------------------------------------------------
__global__ void build_input(const rtx_mlp_sample* s, int n, const float* E, int none, float* X) {
  const int i = blockIdx.x;
  if (i >= n) return;
  rtx_mlp_sample sm = s[i];
  // out-of-range treelets read the "none" row
  if (sm.cur > (uint32_t)none) sm.cur = none;
  for (int h = 0; h < 4; ++h)
    if (sm.hist[h] != RTX_NO_TREELET && sm.hist[h] > (uint32_t)none) sm.hist[h] = RTX_NO_TREELET;
  float* x = X + (size_t)i * kInPad;
  for (int c = threadIdx.x; c < kInPad; c += blockDim.x) {
    float v = 0.f;
    if (c < 32) {
      v = E[(size_t)sm.cur * kEmb + c];
    } else if (c < 40) {
      v = (uint32_t)(c - 32) == sm.octant ? 1.f : 0.f;
    } else if (c == 40) {
      v = sm.depth;
    } else if (c < kIn) {
      const int h = (c - 41) / 32, e = (c - 41) % 32;
      const uint32_t t = sm.hist[h];
      v = E[(size_t)(t == RTX_NO_TREELET ? (uint32_t)none : t) * kEmb + e];
    }
    x[c] = v;
  }
}

__global__ void bias_act(float* Z, int n, int cols, const float* bias, int relu) {
  const size_t total = (size_t)n * cols;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < total; k += (size_t)gridDim.x * blockDim.x) {
    float v = Z[k] + bias[k % cols];
    Z[k] = relu ? (v > 0.f ? v : 0.f) : v;
  }
}

__global__ void relu_mask(float* dH, const float* H, size_t total) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < total; k += (size_t)gridDim.x * blockDim.x)
    if (!(H[k] > 0.f)) dH[k] = 0.f;
}

// column sums of M[n x cols] added into out; blockIdx.y takes rows
// [64 * blockIdx.y, 64 * blockIdx.y + 64)
__global__ void colsum(const float* M, int n, int cols, float* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  const int r0 = blockIdx.y * 64, r1 = min(n, r0 + 64);
  float s = 0.f;
  for (int r = r0; r < r1; ++r) s += M[(size_t)r * cols + c];
  atomicAdd(&out[c], s);
}

__device__ float block_reduce(float v, float* sm, bool is_max) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) {
    float y = __shfl_xor(v, o, 64);
    v = is_max ? fmaxf(v, y) : v + y;
  }
  if (lane == 0) sm[w] = v;
  __syncthreads();
  float r = sm[0];
  for (int k = 1; k < (int)(blockDim.x >> 6); ++k) r = is_max ? fmaxf(r, sm[k]) : r + sm[k];
  __syncthreads();
  return r;
}

// Softmax cross-entropy per row: loss (summed into *loss) and dL = (p - onehot) / n.
__global__ __launch_bounds__(256) void softmax_ce(float* L, int n, int C, const float* cbias,
                                                  const rtx_mlp_sample* s, float* loss, float inv_n, int want_grad) {
  __shared__ float sm[4];
  const int r = blockIdx.x;
  if (r >= n) return;
  float* row = L + (size_t)r * C;
  float mx = -INFINITY;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    row[c] += cbias[c];
    mx = fmaxf(mx, row[c]);
  }
  mx = block_reduce(mx, sm, true);
  float se = 0.f;
  for (int c = threadIdx.x; c < C; c += blockDim.x) se += expf(row[c] - mx);
  // the target logit is read before the barrier inside the next reduction:
  // after it, the thread owning column t may overwrite it with its gradient
  const uint32_t t = s[r].target;
  const float lt = (threadIdx.x == 0 && t < (uint32_t)C) ? row[t] : 0.f;
  se = block_reduce(se, sm, false);
  if (threadIdx.x == 0 && t < (uint32_t)C) atomicAdd(loss, (logf(se) - (lt - mx)) * inv_n);
  if (want_grad)
    for (int c = threadIdx.x; c < C; c += blockDim.x)
      row[c] = (expf(row[c] - mx) / se - ((uint32_t)c == t ? 1.f : 0.f)) * inv_n;
}

__global__ void scatter_embed(const float* dX, const rtx_mlp_sample* s, int n, int none, float* dE) {
  const int i = blockIdx.x;
  if (i >= n) return;
  rtx_mlp_sample sm = s[i];
  if (sm.cur > (uint32_t)none) sm.cur = none;
  for (int h = 0; h < 4; ++h)
    if (sm.hist[h] != RTX_NO_TREELET && sm.hist[h] > (uint32_t)none) sm.hist[h] = RTX_NO_TREELET;
  const float* dx = dX + (size_t)i * kInPad;
  for (int c = threadIdx.x; c < kIn; c += blockDim.x) {
    if (c < 32) {
      atomicAdd(&dE[(size_t)sm.cur * kEmb + c], dx[c]);
    } else if (c >= 41) {
      const int h = (c - 41) / 32, e = (c - 41) % 32;
      const uint32_t t = sm.hist[h];
      atomicAdd(&dE[(size_t)(t == RTX_NO_TREELET ? (uint32_t)none : t) * kEmb + e], dx[c]);
    }
  }
}

__global__ void adam_step(float* p, const float* g, float* m, float* v, size_t count, float lr, float b1, float b2,
                          float eps, float c1, float c2) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count; k += (size_t)gridDim.x * blockDim.x) {
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
  }
}

// Top-4 classes of each row of L (+ bias) and their softmax probabilities,
// one wave per row.
__global__ __launch_bounds__(64) void top4(const float* L, int n, int C, const float* cbias, uint32_t* out,
                                           float* prob) {
  const int r = blockIdx.x;
  if (r >= n) return;
  const int lane = threadIdx.x;
  const float* row = L + (size_t)r * C;
  float bv[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  int bi[4] = {-1, -1, -1, -1};
  for (int c = lane; c < C; c += 64) {
    const float v = row[c] + cbias[c];
    if (v > bv[3]) {
      int k = 3;
      while (k > 0 && v > bv[k - 1]) {
        bv[k] = bv[k - 1];
        bi[k] = bi[k - 1];
        --k;
      }
      bv[k] = v;
      bi[k] = c;
    }
  }
  // merge the 64 lane lists: 4 rounds of a wave argmax (ties to the lower class)
  float topv[4];
  int topi[4];
  for (int round = 0; round < 4; ++round) {
    float v = bv[0];
    int idx = bi[0];
    int src = lane;
    for (int o = 32; o > 0; o >>= 1) {
      const float v2 = __shfl_xor(v, o, 64);
      const int i2 = __shfl_xor(idx, o, 64);
      const int s2 = __shfl_xor(src, o, 64);
      const bool take = i2 >= 0 && (idx < 0 || v2 > v || (v2 == v && i2 < idx));
      if (take) {
        v = v2;
        idx = i2;
        src = s2;
      }
    }
    topv[round] = v;
    topi[round] = idx;
    if (lane == src && idx >= 0) {
      for (int k = 0; k < 3; ++k) {
        bv[k] = bv[k + 1];
        bi[k] = bi[k + 1];
      }
      bv[3] = -INFINITY;
      bi[3] = -1;
    }
  }
  const float mx = topv[0];
  float se = 0.f;
  for (int c = lane; c < C; c += 64) se += expf(row[c] + cbias[c] - mx);
  for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o, 64);
  if (lane < 4) {
    float v = topv[0];
    int idx = topi[0];
    for (int k = 1; k < 4; ++k)
      if (lane == k) {
        v = topv[k];
        idx = topi[k];
      }
    out[(size_t)r * 4 + lane] = idx < 0 ? RTX_NO_TREELET : (uint32_t)idx;
    if (prob) prob[(size_t)r * 4 + lane] = idx < 0 ? 0.f : expf(v - mx) / se;
  }
}

template <bool TA, bool TB, int BN>
hipError_t gemm_launch(hipStream_t s, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
                       int ldc, int accumulate) {
  constexpr int BM = BN == 64 ? 64 : 128;
  const int tiles = ((N + BN - 1) / BN) * ((M + BM - 1) / BM);
  // split K when the output has too few tiles to fill the chip
  int split = 1;
  if (tiles < 256 && K >= 512) split = std::max(1, std::min(K / 256, (512 + tiles - 1) / tiles));
  int k_chunk = (K + split - 1) / split;
  k_chunk = (k_chunk + 31) / 32 * 32;
  split = (K + k_chunk - 1) / k_chunk;
  if (split > 1 && !accumulate) {
    hipLaunchKernelGGL(zero_rows, dim3(256), dim3(256), 0, s, C, M, N, ldc);
  }
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, split);
  hipLaunchKernelGGL((gemm_bf16<TA, TB, BN>), grid, dim3(256), 0, s, M, N, K, k_chunk, A, lda, B, ldb, C, ldc,
                     accumulate);
  return hipGetLastError();
}

template <bool TA, bool TB>
hipError_t gemm(hipStream_t s, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
                int ldc, int accumulate) {
  if (N <= 32) return gemm_launch<TA, TB, 32>(s, M, N, K, A, lda, B, ldb, C, ldc, accumulate);
  return gemm_launch<TA, TB, 64>(s, M, N, K, A, lda, B, ldb, C, ldc, accumulate);
}

}  // namespace

// ---- model object -------------------------------------------------------------
struct rtx_mlp {
  rtx_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  int device = 0;
  uint32_t T = 0, C = 0;  // embedded treelets (row T = "none"), output classes
  // parameter blocks inside one FP32 buffer
  size_t oE, oW1, ob1, oW2, ob2, oW3, ob3, oO, oc, count;
  float *P = nullptr, *G = nullptr, *M1 = nullptr, *V1 = nullptr;
  uint32_t steps = 0;
  // activations of one batch
  uint32_t cap = 0;
  float *X = nullptr, *H1 = nullptr, *H2 = nullptr, *Q = nullptr, *L = nullptr;
  float *dQ = nullptr, *dH2 = nullptr, *dH1 = nullptr, *dX = nullptr, *loss = nullptr;
};

namespace {

void mlp_free_acts(rtx_mlp* m) {
  float** bufs[] = {&m->X, &m->H1, &m->H2, &m->Q, &m->L, &m->dQ, &m->dH2, &m->dH1, &m->dX};
  for (float** b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  m->cap = 0;
}

rtx_status mlp_ensure(rtx_mlp* m, uint32_t b) {
  if (b <= m->cap) return RTX_OK;
  mlp_free_acts(m);
  const size_t sizes[] = {(size_t)b * kInPad, (size_t)b * kHid, (size_t)b * kHid, (size_t)b * kEmb,
                          (size_t)b * m->C,   (size_t)b * kEmb, (size_t)b * kHid, (size_t)b * kHid,
                          (size_t)b * kInPad};
  float** bufs[] = {&m->X, &m->H1, &m->H2, &m->Q, &m->L, &m->dQ, &m->dH2, &m->dH1, &m->dX};
  for (int k = 0; k < 9; ++k)
    if (hipMalloc(bufs[k], sizes[k] * sizeof(float)) != hipSuccess) {
      (void)hipGetLastError();
      mlp_free_acts(m);
      return RTX_ERR_OUT_OF_MEMORY;
    }
  m->cap = b;
  return RTX_OK;
}

#define MK(...)                     \
  do {                              \
    if ((__VA_ARGS__) != hipSuccess) { \
      (void)hipGetLastError();      \
      return RTX_ERR_HIP;           \
    }                               \
  } while (0)

// Forward over n samples; logits stay in m->L (without the output bias).
rtx_status mlp_forward(rtx_mlp* m, const rtx_mlp_sample* s, uint32_t n) {
  hipStream_t st = m->stream;
  float* P = m->P;
  hipLaunchKernelGGL(build_input, dim3(n), dim3(64), 0, st, s, (int)n, P + m->oE, (int)m->T, m->X);
  MK(hipGetLastError());
  MK(gemm<false, false>(st, n, kHid, kInPad, m->X, kInPad, P + m->oW1, kHid, m->H1, kHid, 0));
  hipLaunchKernelGGL(bias_act, dim3(256), dim3(256), 0, st, m->H1, (int)n, kHid, P + m->ob1, 1);
  MK(gemm<false, false>(st, n, kHid, kHid, m->H1, kHid, P + m->oW2, kHid, m->H2, kHid, 0));
  hipLaunchKernelGGL(bias_act, dim3(256), dim3(256), 0, st, m->H2, (int)n, kHid, P + m->ob2, 1);
  MK(gemm<false, false>(st, n, kEmb, kHid, m->H2, kHid, P + m->oW3, kEmb, m->Q, kEmb, 0));
  hipLaunchKernelGGL(bias_act, dim3(64), dim3(256), 0, st, m->Q, (int)n, kEmb, P + m->ob3, 0);
  // logits = Q . O^T  (O is [C x 32])
  MK(gemm<false, true>(st, n, m->C, kEmb, m->Q, kEmb, P + m->oO, kEmb, m->L, m->C, 0));
  MK(hipGetLastError());
  return RTX_OK;
}

// Forward + backward of one batch; gradients are added into m->G.
rtx_status mlp_fwd_bwd(rtx_mlp* m, const rtx_mlp_sample* s, uint32_t n, float inv_n) {
  rtx_status stt = mlp_forward(m, s, n);
  if (stt != RTX_OK) return stt;
  hipStream_t st = m->stream;
  float *P = m->P, *G = m->G;
  hipLaunchKernelGGL(softmax_ce, dim3(n), dim3(256), 0, st, m->L, (int)n, (int)m->C, P + m->oc, s, m->loss, inv_n, 1);
  MK(hipGetLastError());
  // output layer: dO += dL^T . Q ; dc += colsum(dL) ; dQ = dL . O
  MK(gemm<true, false>(st, m->C, kEmb, n, m->L, m->C, m->Q, kEmb, G + m->oO, kEmb, 1));
  hipLaunchKernelGGL(colsum, dim3((m->C + 255) / 256, (n + 63) / 64), dim3(256), 0, st, m->L, (int)n, (int)m->C,
                     G + m->oc);
  MK(gemm<false, false>(st, n, kEmb, m->C, m->L, m->C, P + m->oO, kEmb, m->dQ, kEmb, 0));
  // q = h2 W3 + b3
  MK(gemm<true, false>(st, kHid, kEmb, n, m->H2, kHid, m->dQ, kEmb, G + m->oW3, kEmb, 1));
  hipLaunchKernelGGL(colsum, dim3(1, (n + 63) / 64), dim3(64), 0, st, m->dQ, (int)n, kEmb, G + m->ob3);
  MK(gemm<false, true>(st, n, kHid, kEmb, m->dQ, kEmb, P + m->oW3, kEmb, m->dH2, kHid, 0));
  hipLaunchKernelGGL(relu_mask, dim3(256), dim3(256), 0, st, m->dH2, m->H2, (size_t)n * kHid);
  // h2 = relu(h1 W2 + b2)
  MK(gemm<true, false>(st, kHid, kHid, n, m->H1, kHid, m->dH2, kHid, G + m->oW2, kHid, 1));
  hipLaunchKernelGGL(colsum, dim3(1, (n + 63) / 64), dim3(128), 0, st, m->dH2, (int)n, kHid, G + m->ob2);
  MK(gemm<false, true>(st, n, kHid, kHid, m->dH2, kHid, P + m->oW2, kHid, m->dH1, kHid, 0));
  hipLaunchKernelGGL(relu_mask, dim3(256), dim3(256), 0, st, m->dH1, m->H1, (size_t)n * kHid);
  // h1 = relu(x W1 + b1)
  MK(gemm<true, false>(st, kInPad, kHid, n, m->X, kInPad, m->dH1, kHid, G + m->oW1, kHid, 1));
  hipLaunchKernelGGL(colsum, dim3(1, (n + 63) / 64), dim3(128), 0, st, m->dH1, (int)n, kHid, G + m->ob1);
  MK(gemm<false, true>(st, n, kInPad, kHid, m->dH1, kHid, P + m->oW1, kHid, m->dX, kInPad, 0));
  hipLaunchKernelGGL(scatter_embed, dim3(n), dim3(64), 0, st, m->dX, s, (int)n, (int)m->T, G + m->oE);
  MK(hipGetLastError());
  return RTX_OK;
}

rtx_status mlp_adam(rtx_mlp* m, float lr) {
  m->steps++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)m->steps), c2 = 1.f - powf(b2, (float)m->steps);
  hipLaunchKernelGGL(adam_step, dim3(1024), dim3(256), 0, m->stream, m->P, m->G, m->M1, m->V1, m->count, lr, b1, b2,
                     1e-8f, c1, c2);
  MK(hipGetLastError());
  return RTX_OK;
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace

// Context access (api.cpp).
namespace rtx_internal {
rtx_status ctx_stream(rtx_ctx* c, hipStream_t* s, int* device);
}

extern "C" {

rtx_status rtx_mlp_create(rtx_ctx* c, uint32_t n_treelets, uint32_t n_classes, uint64_t seed, rtx_mlp** out) {
  if (!c || !out || n_treelets == 0 || n_classes == 0) return RTX_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  rtx_mlp* m = new (std::nothrow) rtx_mlp();
  if (!m) return RTX_ERR_OUT_OF_MEMORY;
  m->ctx = c;
  if (rtx_internal::ctx_stream(c, &m->stream, &m->device) != RTX_OK) {
    delete m;
    return RTX_ERR_INVALID_ARGUMENT;
  }
  m->T = n_treelets;
  m->C = n_classes;
  size_t o = 0;
  m->oE = o;  o += (size_t)(n_treelets + 1) * kEmb;
  m->oW1 = o; o += (size_t)kInPad * kHid;
  m->ob1 = o; o += kHid;
  m->oW2 = o; o += (size_t)kHid * kHid;
  m->ob2 = o; o += kHid;
  m->oW3 = o; o += (size_t)kHid * kEmb;
  m->ob3 = o; o += kEmb;
  m->oO = o;  o += (size_t)n_classes * kEmb;
  m->oc = o;  o += n_classes;
  m->count = o;
  (void)hipSetDevice(m->device);
  if (hipMalloc(&m->P, o * 4) != hipSuccess || hipMalloc(&m->G, o * 4) != hipSuccess ||
      hipMalloc(&m->M1, o * 4) != hipSuccess || hipMalloc(&m->V1, o * 4) != hipSuccess ||
      hipMalloc(&m->loss, 4) != hipSuccess) {
    (void)hipGetLastError();
    rtx_mlp_destroy(m);
    return RTX_ERR_OUT_OF_MEMORY;
  }
  // He-style uniform init for weights, small embeddings, zero biases.
  std::vector<float> h(o, 0.f);
  uint64_t s = seed ^ 0x6d6c70ull;
  auto uni = [&](float a) { return a * (2.f * (float)((splitmix(s) >> 40) * (1.0 / 16777216.0)) - 1.f); };
  for (size_t k = m->oE; k < m->oW1; ++k) h[k] = uni(0.5f);
  for (size_t r = 0; r < (size_t)kInPad; ++r)
    for (size_t j = 0; j < (size_t)kHid; ++j) h[m->oW1 + r * kHid + j] = r < (size_t)kIn ? uni(sqrtf(6.f / kIn)) : 0.f;
  for (size_t k = m->oW2; k < m->ob2; ++k) h[k] = uni(sqrtf(6.f / kHid));
  for (size_t k = m->oW3; k < m->ob3; ++k) h[k] = uni(sqrtf(6.f / kHid));
  for (size_t k = m->oO; k < m->oc; ++k) h[k] = uni(sqrtf(3.f / kEmb));
  if (hipMemcpy(m->P, h.data(), o * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->G, 0, o * 4) != hipSuccess || hipMemset(m->M1, 0, o * 4) != hipSuccess ||
      hipMemset(m->V1, 0, o * 4) != hipSuccess) {
    rtx_mlp_destroy(m);
    return RTX_ERR_HIP;
  }
  *out = m;
  return RTX_OK;
}

rtx_status rtx_mlp_destroy(rtx_mlp* m) {
  if (!m) return RTX_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  mlp_free_acts(m);
  float* bufs[] = {m->P, m->G, m->M1, m->V1, m->loss};
  for (float* b : bufs)
    if (b) (void)hipFree(b);
  delete m;
  return RTX_OK;
}

rtx_status rtx_mlp_param_count(const rtx_mlp* m, uint64_t* count) {
  if (!m || !count) return RTX_ERR_INVALID_ARGUMENT;
  *count = m->count;
  return RTX_OK;
}

rtx_status rtx_mlp_get_params(rtx_mlp* m, float* params_host, float* grads_host) {
  if (!m) return RTX_ERR_INVALID_ARGUMENT;
  MK(hipSetDevice(m->device));
  MK(hipStreamSynchronize(m->stream));
  if (params_host) MK(hipMemcpy(params_host, m->P, m->count * 4, hipMemcpyDeviceToHost));
  if (grads_host) MK(hipMemcpy(grads_host, m->G, m->count * 4, hipMemcpyDeviceToHost));
  return RTX_OK;
}

rtx_status rtx_mlp_set_params(rtx_mlp* m, const float* params_host) {
  if (!m || !params_host) return RTX_ERR_INVALID_ARGUMENT;
  MK(hipSetDevice(m->device));
  MK(hipStreamSynchronize(m->stream));
  MK(hipMemcpy(m->P, params_host, m->count * 4, hipMemcpyHostToDevice));
  return RTX_OK;
}

rtx_status rtx_mlp_loss_grad(rtx_mlp* m, const rtx_mlp_sample* samples, uint32_t n, float* loss) {
  if (!m || !samples || n == 0) return RTX_ERR_INVALID_ARGUMENT;
  MK(hipSetDevice(m->device));
  rtx_status st = mlp_ensure(m, n);
  if (st != RTX_OK) return st;
  MK(hipMemsetAsync(m->G, 0, m->count * 4, m->stream));
  MK(hipMemsetAsync(m->loss, 0, 4, m->stream));
  st = mlp_fwd_bwd(m, samples, n, 1.f / (float)n);
  if (st != RTX_OK) return st;
  if (loss) MK(hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
  MK(hipStreamSynchronize(m->stream));
  return RTX_OK;
}

rtx_status rtx_mlp_train(rtx_mlp* m, const rtx_mlp_sample* samples, uint32_t n, uint32_t epochs, uint32_t batch,
                         float lr, float* last_epoch_loss) {
  if (!m || !samples || n == 0 || batch == 0 || epochs == 0) return RTX_ERR_INVALID_ARGUMENT;
  MK(hipSetDevice(m->device));
  batch = std::min(batch, n);
  rtx_status st = mlp_ensure(m, batch);
  if (st != RTX_OK) return st;
  float epoch_loss = 0.f;
  for (uint32_t e = 0; e < epochs; ++e) {
    MK(hipMemsetAsync(m->loss, 0, 4, m->stream));
    uint32_t nb = 0;
    for (uint32_t b0 = 0; b0 < n; b0 += batch) {
      const uint32_t b = std::min(batch, n - b0);
      MK(hipMemsetAsync(m->G, 0, m->count * 4, m->stream));
      st = mlp_fwd_bwd(m, samples + b0, b, 1.f / (float)b);
      if (st != RTX_OK) return st;
      st = mlp_adam(m, lr);
      if (st != RTX_OK) return st;
      ++nb;
    }
    MK(hipMemcpyAsync(&epoch_loss, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
    MK(hipStreamSynchronize(m->stream));
    epoch_loss /= (float)nb;
  }
  if (last_epoch_loss) *last_epoch_loss = epoch_loss;
  return RTX_OK;
}

rtx_status rtx_mlp_predict(rtx_mlp* m, const rtx_mlp_sample* samples, uint32_t n, uint32_t* top4_dev,
                           float* probs_dev) {
  if (!m || !samples || !top4_dev || n == 0) return RTX_ERR_INVALID_ARGUMENT;
  MK(hipSetDevice(m->device));
  const uint32_t chunk = 4096;
  rtx_status st = mlp_ensure(m, std::min(n, chunk));
  if (st != RTX_OK) return st;
  for (uint32_t b0 = 0; b0 < n; b0 += chunk) {
    const uint32_t b = std::min(chunk, n - b0);
    st = mlp_forward(m, samples + b0, b);
    if (st != RTX_OK) return st;
    hipLaunchKernelGGL(top4, dim3(b), dim3(64), 0, m->stream, m->L, (int)b, (int)m->C, m->P + m->oc,
                       top4_dev + (size_t)b0 * 4, probs_dev ? probs_dev + (size_t)b0 * 4 : nullptr);
    MK(hipGetLastError());
  }
  MK(hipStreamSynchronize(m->stream));
  return RTX_OK;
}

rtx_status rtx_mlp_gemm(rtx_ctx* c, int M, int N, int K, int trans_a, int trans_b, const float* A_dev,
                              const float* B_dev, float* C_dev) {
  hipStream_t s;
  int dev;
  if (!c || rtx_internal::ctx_stream(c, &s, &dev) != RTX_OK || M <= 0 || N <= 0 || K <= 0) return RTX_ERR_INVALID_ARGUMENT;
  MK(hipSetDevice(dev));
  const int lda = trans_a ? M : K, ldb = trans_b ? K : N;
  hipError_t e;
  if (!trans_a && !trans_b) e = gemm<false, false>(s, M, N, K, A_dev, lda, B_dev, ldb, C_dev, N, 0);
  else if (!trans_a && trans_b) e = gemm<false, true>(s, M, N, K, A_dev, lda, B_dev, ldb, C_dev, N, 0);
  else if (trans_a && !trans_b) e = gemm<true, false>(s, M, N, K, A_dev, lda, B_dev, ldb, C_dev, N, 0);
  else e = gemm<true, true>(s, M, N, K, A_dev, lda, B_dev, ldb, C_dev, N, 0);
  MK(e);
  MK(hipStreamSynchronize(s));
  return RTX_OK;
}

}  // extern "C"
