// Forward + backward of the hop-proposal network (16 -> 64 -> 64 -> 1, ReLU,
// mean binary cross-entropy with logits) for training on recorded moves.  No
// reference counterpart.  One workgroup per tile of 64 samples: the 64x64x64
// products (layer-2 forward, dH1 = dH2 W2, dW2 = dH2^T H1) run on the matrix
// cores as v_mfma_f32_16x16x16bf16 (BF16 operands, FP32 accumulation), one
// wavefront per 16-row tile; the thin layers (16 inputs, 1 output) are FP32 FMA.
// Per-tile partial gradients are summed in tile order by a second kernel, so the
// result does not depend on scheduling.
#include "common.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

#define NP_MLP (16 * 64 + 64 + 64 * 64 + 64 + 64 + 1)
#define OFF_W1 0
#define OFF_B1 (16 * 64)
#define OFF_W2 (OFF_B1 + 64)
#define OFF_B2 (OFF_W2 + 64 * 64)
#define OFF_W3 (OFF_B2 + 64)
#define OFF_B3 (OFF_W3 + 64)

__device__ __forceinline__ short to_bf16(float x) {  // round to nearest even
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// LDS row pitch of 72 bf16 (64 + 8) keeps the 4-element column gathers off one bank
#define PB 72
struct MlpShared {
  float x[MLP_TILE][16];
  float w1[64 * 16];
  float b1[64], b2[64], w3[64];
  short w2[64 * PB];      // W2[j][i]
  float a1[MLP_TILE][65]; // layer-1 pre-activation
  short h1[MLP_TILE * PB];
  float a2[MLP_TILE][65]; // layer-2 pre-activation
  short dh2[MLP_TILE * PB];
  float dz[MLP_TILE];
  float lsum[MLP_TILE];
};

__global__ __launch_bounds__(256) void k_mlp_tile(const float* __restrict__ w, const float* __restrict__ X,
                                                   const float* __restrict__ Y, int n, float* part) {
  __shared__ MlpShared S;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int s0 = blockIdx.x * MLP_TILE;
  const int cnt = min(MLP_TILE, n - s0);
  for (int i = tid; i < 64 * 16; i += 256) S.w1[i] = w[OFF_W1 + i];
  for (int i = tid; i < 64 * 64; i += 256) S.w2[(i >> 6) * PB + (i & 63)] = to_bf16(w[OFF_W2 + i]);
  if (tid < 64) {
    S.b1[tid] = w[OFF_B1 + tid];
    S.b2[tid] = w[OFF_B2 + tid];
    S.w3[tid] = w[OFF_W3 + tid];
  }
  for (int i = tid; i < MLP_TILE * 16; i += 256) {
    int s = i >> 4;
    S.x[s][i & 15] = (s < cnt) ? X[(size_t)(s0 + s) * 16 + (i & 15)] : 0.f;
  }
  __syncthreads();
  // layer 1 (FP32): a1[s][j] = b1[j] + sum_i W1[j][i] x[s][i]
  for (int e = tid; e < MLP_TILE * 64; e += 256) {
    int s = e >> 6, j = e & 63;
    float a = S.b1[j];
#pragma unroll
    for (int i = 0; i < 16; ++i) a = fmaf(S.w1[j * 16 + i], S.x[s][i], a);
    S.a1[s][j] = a;
    S.h1[s * PB + j] = to_bf16(fmaxf(a, 0.f));
  }
  __syncthreads();
  // layer 2 on the matrix cores: a2 = h1 W2^T + b2, wave wv owns samples 16wv..16wv+15
  const int r16 = lane & 15, q4 = lane >> 4;
  for (int nt = 0; nt < 4; ++nt) {
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      bf16x4 a, b;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        a[t] = S.h1[(16 * wv + r16) * PB + 16 * kt + 4 * q4 + t];
        b[t] = S.w2[(16 * nt + r16) * PB + 16 * kt + 4 * q4 + t];  // B[k][n] = W2[n][k]
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      int s = 16 * wv + 4 * q4 + t, j = 16 * nt + r16;
      S.a2[s][j] = acc[t] + S.b2[j];
    }
  }
  __syncthreads();
  // output, loss, dz
  if (tid < MLP_TILE) {
    const int s = tid;
    float z = 0.f;
    for (int j = 0; j < 64; ++j) z = fmaf(fmaxf(S.a2[s][j], 0.f), S.w3[j], z);
    z += w[OFF_B3];
    float l = 0.f, d = 0.f;
    if (s < cnt) {
      const float y = Y[s0 + s];
      l = fmaxf(z, 0.f) - z * y + log1pf(expf(-fabsf(z)));
      d = (1.f / (1.f + expf(-z)) - y) / (float)n;
    }
    S.dz[s] = d;
    S.lsum[s] = l;
  }
  __syncthreads();
  float* P = part + (size_t)blockIdx.x * (NP_MLP + 1);
  // dH2 = dz W3^T (.) [a2 > 0]  (BF16 for the matrix cores)
  for (int e = tid; e < MLP_TILE * 64; e += 256) {
    int s = e >> 6, j = e & 63;
    S.dh2[s * PB + j] = to_bf16(S.a2[s][j] > 0.f ? S.dz[s] * S.w3[j] : 0.f);
  }
  // dW3, b3, b2 partials (FP32, over the tile's samples)
  if (tid < 64) {
    const int j = tid;
    float g3 = 0.f, g2 = 0.f;
    for (int s = 0; s < MLP_TILE; ++s) {
      float h = fmaxf(S.a2[s][j], 0.f);
      g3 = fmaf(S.dz[s], h, g3);
      g2 += S.a2[s][j] > 0.f ? S.dz[s] * S.w3[j] : 0.f;
    }
    P[OFF_W3 + j] = g3;
    P[OFF_B2 + j] = g2;
  } else if (tid == 64) {
    float g = 0.f, l = 0.f;
    for (int s = 0; s < MLP_TILE; ++s) {
      g += S.dz[s];
      l += S.lsum[s];
    }
    P[OFF_B3] = g;
    P[NP_MLP] = l;
  }
  __syncthreads();
  // dW2[j][i] = sum_s dH2[s][j] h1[s][i]: wave wv owns rows j = 16wv..16wv+15
  for (int nt = 0; nt < 4; ++nt) {
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      bf16x4 a, b;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        int s = 16 * kt + 4 * q4 + t;
        a[t] = S.dh2[s * PB + 16 * wv + r16];  // A[m=j][k=s]
        b[t] = S.h1[s * PB + 16 * nt + r16];   // B[k=s][n=i]
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) P[OFF_W2 + (16 * wv + 4 * q4 + t) * 64 + 16 * nt + r16] = acc[t];
  }
  // dH1 = (dH2 W2) (.) [a1 > 0]: wave wv owns samples 16wv..; reuse a2 as dH1 storage
  for (int nt = 0; nt < 4; ++nt) {
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      bf16x4 a, b;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        a[t] = S.dh2[(16 * wv + r16) * PB + 16 * kt + 4 * q4 + t];  // A[m=s][k=j]
        b[t] = S.w2[(16 * kt + 4 * q4 + t) * PB + 16 * nt + r16];   // B[k=j][n=i] = W2[j][i]
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      int s = 16 * wv + 4 * q4 + t, i = 16 * nt + r16;
      S.a2[s][i] = S.a1[s][i] > 0.f ? acc[t] : 0.f;
    }
  }
  __syncthreads();
  // dW1[j][i] = sum_s dH1[s][j] x[s][i]; db1[j] = sum_s dH1[s][j]
  for (int e = tid; e < 64 * 16; e += 256) {
    int j = e >> 4, i = e & 15;
    float g = 0.f;
    for (int s = 0; s < MLP_TILE; ++s) g = fmaf(S.a2[s][j], S.x[s][i], g);
    P[OFF_W1 + e] = g;
  }
  if (tid < 64) {
    float g = 0.f;
    for (int s = 0; s < MLP_TILE; ++s) g += S.a2[s][tid];
    P[OFF_B1 + tid] = g;
  }
}

// out[p] = sum over tiles of part[tile][p] (tile order); out[NP] = mean loss
__global__ void k_mlp_reduce(const float* part, int ntile, int n, float* out) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > NP_MLP) return;
  float g = 0.f;
  for (int b = 0; b < ntile; ++b) g += part[(size_t)b * (NP_MLP + 1) + p];
  out[p] = (p == NP_MLP) ? g / (float)n : g;
}

void launch_mlp_loss_grad(hipStream_t st, const float* w, const float* X, const float* Y, int n, float* part,
                          float* out) {
  if (n <= 0) return;
  const int nt = (n + MLP_TILE - 1) / MLP_TILE;
  k_mlp_tile<<<nt, 256, 0, st>>>(w, X, Y, n, part);
  k_mlp_reduce<<<(NP_MLP + 1 + 255) / 256, 256, 0, st>>>(part, nt, n, out);
}
