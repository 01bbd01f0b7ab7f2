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
// [kColsumRows * blockIdx.y, kColsumRows * blockIdx.y + kColsumRows)
constexpr int kColsumRows = 8;
__global__ void colsum(const float* M, int n, int cols, float* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  const int r0 = blockIdx.y * kColsumRows, r1 = min(n, r0 + kColsumRows);
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

// dE[row of slot h of sample i][e] += dX[i][column of slot h, e] for the five
// embedding slots (cur, hist 0-3).  One wave takes 64 samples of one (slot,
// column) pair and adds the values of lanes that name the same row with one
// atomic: the ancestors in the history are shared by most samples, and
// per-sample atomics on those rows serialise.  grid.x: 64-sample groups,
// grid.y: 5 slots x 32 columns.
__global__ __launch_bounds__(64) void scatter_embed(const float* dX, const rtx_mlp_sample* s, int n, int none,
                                                   float* dE) {
  const int lane = threadIdx.x;
  const int i = blockIdx.x * 64 + lane;
  const int h = blockIdx.y / kEmb, e = blockIdx.y % kEmb;
  uint32_t row = 0xffffffffu;
  float v = 0.f;
  if (i < n) {
    const rtx_mlp_sample sm = s[i];
    uint32_t t;
    int col;
    if (h == 0) {
      t = sm.cur > (uint32_t)none ? (uint32_t)none : sm.cur;  // out-of-range treelets read the "none" row
      col = e;
    } else {
      t = sm.hist[h - 1];
      if (t == RTX_NO_TREELET || t > (uint32_t)none) t = (uint32_t)none;
      col = 41 + (h - 1) * kEmb + e;
    }
    row = t;
    v = dX[(size_t)i * kInPad + col];
  }
  uint64_t left = __ballot(i < n);
  while (left) {
    const int lead = __ffsll((long long)left) - 1;
    const uint32_t key = __shfl(row, lead, 64);
    const bool mine = (left >> lane) & 1ull && row == key;
    float x = mine ? v : 0.f;
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == lead) atomicAdd(&dE[(size_t)key * kEmb + e], x);
    left &= ~__ballot(mine);
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

// ---- fused output layer ----------------------------------------------------------
// The logits L = Q O^T + o (Q: [n x 32] queries, O: [C x 32] class vectors, o the
// class biases) exist only as 16x16 MFMA tiles in registers.  ce_stats keeps each
// row's running maximum and exp-sum over a chunk of classes; ce_finish turns the
// chunks into the row's log-sum-exp and the loss; two gradient passes recompute
// the tiles to form dL = (softmax - onehot) / n: ce_grad_o per 64 classes and
// row split (dO = dL^T Q and the bias gradient) and ce_grad_q per 64 rows and
// class split (dQ = dL O), each summing its splits atomically.  The sum over k inside one MFMA does not depend on which k-slot
// holds which term, so a dL tile left in the result layout (lane = column, rows
// 4g..4g+3 of lane group g = lane / 16) feeds the next MFMA as it is: k-slot
// 8g + j pairs tile row 4g + j (j < 4) of the first 16-row tile with the same
// k-slot of the other operand, and 4g + j - 4 of the second tile for j >= 4.
constexpr int kCeChunk = 256;  // classes per ce_stats block
constexpr int kCeSplit = 32;   // class splits of ce_grad_q
constexpr int kCeRowSplit = 4; // row splits of ce_grad_o
constexpr int kCeRows = 512;   // rows per LDS stage of ce_grad_o
constexpr int kCeCls = 256;    // classes per LDS stage of ce_grad_q

// Eight FP32 values rounded to BF16 (round to nearest even, as the MFMA operand
// loads of gemm_bf16), packed into 16 bytes.
__device__ __forceinline__ uint4 pack_bf16x8(const float* __restrict__ p) {
  const float4 a = *reinterpret_cast<const float4*>(p);
  const float4 b = *reinterpret_cast<const float4*>(p + 4);
  bf16x8 v;
  v[0] = (__bf16)a.x; v[1] = (__bf16)a.y; v[2] = (__bf16)a.z; v[3] = (__bf16)a.w;
  v[4] = (__bf16)b.x; v[5] = (__bf16)b.y; v[6] = (__bf16)b.z; v[7] = (__bf16)b.w;
  return __builtin_bit_cast(uint4, v);
}

// Row r of an [rows x 32] FP32 matrix, elements k0..k0+7, as BF16 (zero past `rows`).
__device__ __forceinline__ bf16x8 ld_row8(const float* __restrict__ M, int rows, int r, int k0) {
  if (r < rows) return __builtin_bit_cast(bf16x8, pack_bf16x8(M + (size_t)r * kEmb + k0));
  return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
}

// Rows [r0, r0 + count) of an [rows x 32] FP32 matrix into LDS as BF16: row i
// at S[4 * i .. 4 * i + 3] (16 bytes = 8 values each).  Every thread of the block calls it.
__device__ __forceinline__ void stage_rows(uint4* S, const float* __restrict__ M, int r0, int count) {
  for (int u = threadIdx.x; u < count * 4; u += blockDim.x) S[u] = pack_bf16x8(M + (size_t)(r0 + u / 4) * kEmb + (u % 4) * 8);
}

__device__ __forceinline__ bf16x8 lds_row8(const uint4* S, int i, int g, bool valid) {
  return __builtin_bit_cast(bf16x8, valid ? S[4 * i + g] : make_uint4(0, 0, 0, 0));
}

__device__ __forceinline__ __bf16 lds_elem(const uint4* S, int i, int col) {
  return reinterpret_cast<const __bf16*>(S)[i * kEmb + col];
}

// The tile row paired with k-slot 8g + j (see above).
__device__ __forceinline__ int kslot_row(int g, int j) { return j < 4 ? 4 * g + j : 16 + 4 * g + j - 4; }

// (m, s) <- the log-sum-exp pair of the union of two sets: max and exp-sum relative to it.
__device__ __forceinline__ void lse_merge(float& m, float& s, float m2, float s2) {
  const float mx = fmaxf(m, m2);
  if (mx == -INFINITY) return;
  s = s * expf(m - mx) + s2 * expf(m2 - mx);
  m = mx;
}

// One block: 64 rows (a wave per 16) x kCeChunk classes (staged in LDS).
// pm/ps[r * nchunk + chunk] get the row's maximum logit and its exp-sum over the
// chunk; tl[r] the target's logit.
__global__ __launch_bounds__(256) void ce_stats(const float* __restrict__ Q, int n, const float* __restrict__ O,
                                                const float* __restrict__ cb, int C,
                                                const rtx_mlp_sample* __restrict__ s, int nchunk, float* pm,
                                                float* ps, float* tl) {
  __shared__ uint4 Os[kCeChunk * 4];
  __shared__ float cbs[kCeChunk];
  const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
  const int r0 = blockIdx.x * 64 + (threadIdx.x >> 6) * 16;
  const int chunk = blockIdx.y;
  const int c_lo = chunk * kCeChunk, nc = min(kCeChunk, C - c_lo);
  stage_rows(Os, O, c_lo, nc);
  for (int c = threadIdx.x; c < nc; c += blockDim.x) cbs[c] = cb[c_lo + c];
  __syncthreads();
  const bf16x8 a = ld_row8(Q, n, r0 + c16, 8 * g);
  float m[4], sm[4];
  uint32_t t[4];
  for (int i = 0; i < 4; ++i) {
    const int r = r0 + 4 * g + i;
    m[i] = -INFINITY;
    sm[i] = 0.f;
    t[i] = r < n ? s[r].target : 0xffffffffu;
  }
  for (int ct = 0; ct < nc; ct += 16) {
    const int cl = ct + c16;
    const bf16x8 b = lds_row8(Os, cl, g, cl < nc);
    const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    if (cl < nc) {
      const float bias = cbs[cl];
      for (int i = 0; i < 4; ++i) {
        const float x = acc[i] + bias;
        const float mx = fmaxf(m[i], x);
        sm[i] = sm[i] * expf(m[i] - mx) + expf(x - mx);
        m[i] = mx;
        if ((uint32_t)(c_lo + cl) == t[i]) tl[r0 + 4 * g + i] = x;
      }
    }
  }
  // the 16 lanes of a group hold the same rows
  for (int o = 1; o < 16; o <<= 1)
    for (int i = 0; i < 4; ++i) {
      const float m2 = __shfl_xor(m[i], o, 64), s2 = __shfl_xor(sm[i], o, 64);
      lse_merge(m[i], sm[i], m2, s2);
    }
  if (c16 == 0)
    for (int i = 0; i < 4; ++i) {
      const int r = r0 + 4 * g + i;
      if (r < n) {
        pm[(size_t)r * nchunk + chunk] = m[i];
        ps[(size_t)r * nchunk + chunk] = sm[i];
      }
    }
}

// Per row: log-sum-exp over the chunks, the loss (one atomic per block of 16
// rows into *loss), and dQ cleared.  A wave takes 4 rows in turn.
__global__ __launch_bounds__(256) void ce_finish(const float* pm, const float* ps, int nchunk, const float* tl,
                                                 const rtx_mlp_sample* s, int n, int C, float inv_n, float* lse,
                                                 float* loss, float* dQ) {
  __shared__ float part[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float acc = 0.f;
  for (int k = 0; k < 4; ++k) {
    const int r = blockIdx.x * 16 + w * 4 + k;
    if (r >= n) break;  // wave-uniform
    float m = -INFINITY, sm = 0.f;
    for (int c = lane; c < nchunk; c += 64) lse_merge(m, sm, pm[(size_t)r * nchunk + c], ps[(size_t)r * nchunk + c]);
    for (int o = 32; o > 0; o >>= 1) {
      const float m2 = __shfl_xor(m, o, 64), s2 = __shfl_xor(sm, o, 64);
      lse_merge(m, sm, m2, s2);
    }
    if (lane < kEmb) dQ[(size_t)r * kEmb + lane] = 0.f;
    const float l = m + logf(sm);
    if (lane == 0) {
      lse[r] = l;
      if (s[r].target < (uint32_t)C) acc += (l - tl[r]) * inv_n;
    }
  }
  if (lane == 0) part[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, part[0] + part[1] + part[2] + part[3]);
}

// dL of one logit: (softmax - onehot) / n.
__device__ __forceinline__ float ce_dl(float logit, float bias, float lse, uint32_t target, int cls, float inv_n) {
  return (expf(logit + bias - lse) - (target == (uint32_t)cls ? 1.f : 0.f)) * inv_n;
}

// One block: 64 classes (a wave per 16) against the rows [y * rper, y * rper + rper),
// staged in LDS kCeRows at a time.  gO[c] += sum_r dL[r, c] Q[r], gc[c] += sum_r dL[r, c]
// (atomically: the row splits of a class add up).
__global__ __launch_bounds__(256) void ce_grad_o(const float* __restrict__ Q, int n, const float* __restrict__ O,
                                                 const float* __restrict__ cb, int C,
                                                 const rtx_mlp_sample* __restrict__ s, const float* __restrict__ lse,
                                                 float inv_n, int rper, float* gO, float* gc) {
  __shared__ uint4 Qs[kCeRows * 4];
  __shared__ float ls[kCeRows];
  __shared__ uint32_t ts[kCeRows];
  const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
  const int c0 = blockIdx.x * 64 + (threadIdx.x >> 6) * 16;
  const int cls = c0 + c16;
  const bool valid = cls < C;
  const bf16x8 b = ld_row8(O, C, cls, 8 * g);
  const float bias = valid ? cb[cls] : 0.f;
  f32x4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
  float dc = 0.f;
  const int rbeg = blockIdx.y * rper, rend = min(n, rbeg + rper);
  for (int rc = rbeg; rc < rend; rc += kCeRows) {
    const int nr = min(kCeRows, rend - rc);
    __syncthreads();  // the previous stage is read
    stage_rows(Qs, Q, rc, nr);
    for (int i = threadIdx.x; i < nr; i += blockDim.x) {
      ls[i] = lse[rc + i];
      ts[i] = s[rc + i].target;
    }
    __syncthreads();
    for (int r0 = 0; r0 < nr; r0 += 32) {
      bf16x8 a2;
      for (int tt = 0; tt < 2; ++tt) {
        const int rt = r0 + 16 * tt;
        const bf16x8 a = lds_row8(Qs, rt + c16, g, rt + c16 < nr);
        const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        for (int i = 0; i < 4; ++i) {
          const int r = rt + 4 * g + i;
          const float v = (valid && r < nr) ? ce_dl(acc[i], bias, ls[r], ts[r], cls, inv_n) : 0.f;
          dc += v;
          a2[4 * tt + i] = (__bf16)v;
        }
      }
      // dO tile [16 classes x 16 emb] += dL^T (classes x 32 rows) . Q (32 rows x 16 emb)
      bf16x8 q0, q1;
      for (int j = 0; j < 8; ++j) {
        const int r = r0 + kslot_row(g, j);
        const bool in = r < nr;
        q0[j] = in ? lds_elem(Qs, r, c16) : (__bf16)0.f;
        q1[j] = in ? lds_elem(Qs, r, 16 + c16) : (__bf16)0.f;
      }
      d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, q0, d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, q1, d1, 0, 0, 0);
    }
  }
  dc += __shfl_xor(dc, 16, 64);
  dc += __shfl_xor(dc, 32, 64);
  if (g == 0 && valid) atomicAdd(&gc[cls], dc);
  // result layout: emb column c16 (+16), classes c0 + 4g + i
  for (int i = 0; i < 4; ++i) {
    const int c2 = c0 + 4 * g + i;
    if (c2 < C) {
      atomicAdd(&gO[(size_t)c2 * kEmb + c16], d0[i]);
      atomicAdd(&gO[(size_t)c2 * kEmb + 16 + c16], d1[i]);
    }
  }
}

// One block: 64 rows (a wave per 16) against the classes [y * cper, y * cper + cper),
// staged in LDS kCeCls at a time.  dQ[r] += sum_c dL[r, c] O[c] (dQ cleared by ce_finish).
__global__ __launch_bounds__(256) void ce_grad_q(const float* __restrict__ Q, int n, const float* __restrict__ O,
                                                 const float* __restrict__ cb, int C,
                                                 const rtx_mlp_sample* __restrict__ s, const float* __restrict__ lse,
                                                 float inv_n, int cper, float* dQ) {
  __shared__ uint4 Os[kCeCls * 4];
  __shared__ float cbs[kCeCls];
  const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
  const int row = blockIdx.x * 64 + (threadIdx.x >> 6) * 16 + c16;
  const bool vr = row < n;
  const bf16x8 b = ld_row8(Q, n, row, 8 * g);  // the row's query: column c16 of the transposed logits
  const float l = vr ? lse[row] : 0.f;
  const uint32_t t = vr ? s[row].target : 0xffffffffu;
  const int cbeg = blockIdx.y * cper, cend = min(C, cbeg + cper);
  f32x4 q0 = {0.f, 0.f, 0.f, 0.f}, q1 = {0.f, 0.f, 0.f, 0.f};
  for (int cc = cbeg; cc < cend; cc += kCeCls) {
    const int nc = min(kCeCls, cend - cc);
    __syncthreads();  // the previous stage is read
    stage_rows(Os, O, cc, nc);
    for (int i = threadIdx.x; i < nc; i += blockDim.x) cbs[i] = cb[cc + i];
    __syncthreads();
    for (int cb0 = 0; cb0 < nc; cb0 += 32) {
      bf16x8 b2;
      for (int tt = 0; tt < 2; ++tt) {
        const int ct = cb0 + 16 * tt;
        // logits^T tile: classes ct + 4g + i (rows of the result) against this lane's row
        const bf16x8 a = lds_row8(Os, ct + c16, g, ct + c16 < nc);
        const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        for (int i = 0; i < 4; ++i) {
          const int cl = ct + 4 * g + i;
          const float v = (vr && cl < nc) ? ce_dl(acc[i], cbs[cl], l, t, cc + cl, inv_n) : 0.f;
          b2[4 * tt + i] = (__bf16)v;
        }
      }
      // dQ^T tile [16 emb x 16 rows] += O^T (emb x 32 classes) . dL^T (32 classes x rows)
      bf16x8 o0, o1;
      for (int j = 0; j < 8; ++j) {
        const int cl = cb0 + kslot_row(g, j);
        const bool in = cl < nc;
        o0[j] = in ? lds_elem(Os, cl, c16) : (__bf16)0.f;
        o1[j] = in ? lds_elem(Os, cl, 16 + c16) : (__bf16)0.f;
      }
      q0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o0, b2, q0, 0, 0, 0);
      q1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o1, b2, q1, 0, 0, 0);
    }
  }
  // result layout: row column c16, emb 4g + i (+16)
  if (vr)
    for (int i = 0; i < 4; ++i) {
      atomicAdd(&dQ[(size_t)row * kEmb + 4 * g + i], q0[i]);
      atomicAdd(&dQ[(size_t)row * kEmb + 16 + 4 * g + i], q1[i]);
    }
}

template <bool TA, bool TB, int BN>
hipError_t gemm_launch(hipStream_t s, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
                       int ldc, int accumulate) {
  constexpr int BM = BN == 64 ? 64 : 128;
  const int tiles = ((N + BN - 1) / BN) * ((M + BM - 1) / BM);
  // split K when the output has too few tiles to fill the chip (the weight
  // gradients: a few tiles, K = the batch), down to 64-deep chunks (32 for
  // outputs of under 8 tiles)
  int split = 1;
  if (tiles < 256 && K >= 512) split = std::max(1, std::min(K / (tiles < 8 ? 32 : 64), (512 + tiles - 1) / tiles));
  // the batch-sized activations (a few dozen tiles, K 128-192): 64-deep chunks
  else if (tiles < 64 && K >= 128) split = std::max(1, std::min(K / 64, (128 + tiles - 1) / tiles));
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
  // fused output layer: per-row chunk maxima and exp-sums, target logits, log-sum-exp
  float *pm = nullptr, *ps = nullptr, *tl = nullptr, *lse = nullptr;
};

namespace {

int ce_chunks(const rtx_mlp* m) { return (int)((m->C + kCeChunk - 1) / kCeChunk); }

void mlp_free_acts(rtx_mlp* m) {
  float** bufs[] = {&m->X, &m->H1, &m->H2, &m->Q, &m->L, &m->dQ, &m->dH2, &m->dH1, &m->dX,
                    &m->pm, &m->ps, &m->tl, &m->lse};
  for (float** b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  m->cap = 0;
}

rtx_status mlp_ensure(rtx_mlp* m, uint32_t b) {
  if (b <= m->cap) return RTX_OK;
  mlp_free_acts(m);
  const size_t nch = (size_t)ce_chunks(m);
  const size_t sizes[] = {(size_t)b * kInPad, (size_t)b * kHid, (size_t)b * kHid, (size_t)b * kEmb,
                          (size_t)b * m->C,   (size_t)b * kEmb, (size_t)b * kHid, (size_t)b * kHid,
                          (size_t)b * kInPad, (size_t)b * nch,  (size_t)b * nch,  (size_t)b,
                          (size_t)b};
  float** bufs[] = {&m->X, &m->H1, &m->H2, &m->Q, &m->L, &m->dQ, &m->dH2, &m->dH1, &m->dX,
                    &m->pm, &m->ps, &m->tl, &m->lse};
  for (int k = 0; k < 13; ++k)
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

// Forward over n samples; logits stay in m->L (without the output bias) unless
// `logits` is false (the fused output layer computes them itself).
rtx_status mlp_forward(rtx_mlp* m, const rtx_mlp_sample* s, uint32_t n, bool logits = true) {
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
  if (logits) MK(gemm<false, true>(st, n, m->C, kEmb, m->Q, kEmb, P + m->oO, kEmb, m->L, m->C, 0));
  MK(hipGetLastError());
  return RTX_OK;
}

// Forward + backward of one batch; gradients are added into m->G.
rtx_status mlp_fwd_bwd(rtx_mlp* m, const rtx_mlp_sample* s, uint32_t n, float inv_n) {
  rtx_status stt = mlp_forward(m, s, n, false);
  if (stt != RTX_OK) return stt;
  hipStream_t st = m->stream;
  float *P = m->P, *G = m->G;
  // output layer, fused: softmax cross-entropy, dO += dL^T . Q, dc += colsum(dL),
  // dQ = dL . O, the logits never in memory
  const int C = (int)m->C, nch = ce_chunks(m);
  const int row_blocks = (int)((n + 63) / 64);
  hipLaunchKernelGGL(ce_stats, dim3(row_blocks, nch), dim3(256), 0, st, m->Q, (int)n, P + m->oO, P + m->oc, C, s, nch,
                     m->pm, m->ps, m->tl);
  hipLaunchKernelGGL(ce_finish, dim3((n + 15) / 16), dim3(256), 0, st, m->pm, m->ps, nch, m->tl, s, (int)n, C, inv_n,
                     m->lse, m->loss, m->dQ);
  const int rsplits = std::max(1, std::min(kCeRowSplit, (int)((n + 31) / 32)));
  const int rper = (int)((n + rsplits - 1) / rsplits + 31) / 32 * 32;
  hipLaunchKernelGGL(ce_grad_o, dim3((C + 63) / 64, rsplits), dim3(256), 0, st, m->Q, (int)n, P + m->oO, P + m->oc, C,
                     s, m->lse, inv_n, rper, G + m->oO, G + m->oc);
  const int splits = std::max(1, std::min(kCeSplit, (C + 31) / 32));
  const int cper = ((C + splits - 1) / splits + 31) / 32 * 32;
  hipLaunchKernelGGL(ce_grad_q, dim3(row_blocks, splits), dim3(256), 0, st, m->Q, (int)n, P + m->oO, P + m->oc, C, s,
                     m->lse, inv_n, cper, m->dQ);
  MK(hipGetLastError());
  // q = h2 W3 + b3
  MK(gemm<true, false>(st, kHid, kEmb, n, m->H2, kHid, m->dQ, kEmb, G + m->oW3, kEmb, 1));
  hipLaunchKernelGGL(colsum, dim3(1, (n + kColsumRows - 1) / kColsumRows), dim3(64), 0, st, m->dQ, (int)n, kEmb, G + m->ob3);
  MK(gemm<false, true>(st, n, kHid, kEmb, m->dQ, kEmb, P + m->oW3, kEmb, m->dH2, kHid, 0));
  hipLaunchKernelGGL(relu_mask, dim3(256), dim3(256), 0, st, m->dH2, m->H2, (size_t)n * kHid);
  // h2 = relu(h1 W2 + b2)
  MK(gemm<true, false>(st, kHid, kHid, n, m->H1, kHid, m->dH2, kHid, G + m->oW2, kHid, 1));
  hipLaunchKernelGGL(colsum, dim3(1, (n + kColsumRows - 1) / kColsumRows), dim3(128), 0, st, m->dH2, (int)n, kHid, G + m->ob2);
  MK(gemm<false, true>(st, n, kHid, kHid, m->dH2, kHid, P + m->oW2, kHid, m->dH1, kHid, 0));
  hipLaunchKernelGGL(relu_mask, dim3(256), dim3(256), 0, st, m->dH1, m->H1, (size_t)n * kHid);
  // h1 = relu(x W1 + b1)
  MK(gemm<true, false>(st, kInPad, kHid, n, m->X, kInPad, m->dH1, kHid, G + m->oW1, kHid, 1));
  hipLaunchKernelGGL(colsum, dim3(1, (n + kColsumRows - 1) / kColsumRows), dim3(128), 0, st, m->dH1, (int)n, kHid, G + m->ob1);
  MK(gemm<false, true>(st, n, kInPad, kHid, m->dH1, kHid, P + m->oW1, kHid, m->dX, kInPad, 0));
  hipLaunchKernelGGL(scatter_embed, dim3((n + 63) / 64, 5 * kEmb), dim3(64), 0, st, m->dX, s, (int)n, (int)m->T,
                     G + m->oE);
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
