// e13_mlp.hip -- the subtree cost model: a 32 -> 64 -> 64 -> 2 perceptron
// (ReLU hidden layers, linear outputs) with hand-written gfx950 kernels for
// inference, the gradient of the mean squared error, and Adam.
//
// Parameter layout (fp32, 6402 values), shared with the host code:
//   W1[64][32] at 0, b1[64] at 2048, W2[64][64] at 2112, b2[64] at 6208,
//   W3[2][64] at 6272, b3[2] at 6400.
// Loss over a batch of n samples: L = (1/n) sum_s sum_o (y[s][o] - t[s][o])^2.
//
// One workgroup (256 threads) takes a tile of 32 samples.  The weights sit in
// LDS with padded rows (33 and 65 floats), so both the row-wise reads of the
// forward pass and the column-wise reads of the backward pass fall in distinct
// banks.  Gradients are written per tile and summed by a second kernel in tile
// order, so a gradient is the same bit pattern on every run.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int NI = 32, NH = 64, NO = 2, TILE = 32;
constexpr int OW1 = 0, OB1 = 2048, OW2 = 2112, OB2 = 6208, OW3 = 6272, OB3 = 6400, NP = 6402;
constexpr int S1 = NI + 1, S2 = NH + 1; // padded LDS row strides
// LDS words
constexpr int L_W1 = 0, L_B1 = L_W1 + NH * S1, L_W2 = L_B1 + NH, L_B2 = L_W2 + NH * S2, L_W3 = L_B2 + NH,
              L_B3 = L_W3 + NO * S2, L_X = L_B3 + 2, L_H1 = L_X + TILE * NI, L_H2 = L_H1 + TILE * NH,
              L_D1 = L_H2 + TILE * NH, L_D2 = L_D1 + TILE * NH, L_Y = L_D2 + TILE * NH, L_DY = L_Y + TILE * NO,
              L_END = L_DY + TILE * NO;
static_assert(L_END * 4 <= 65536, "MLP tile LDS");

} // namespace

// MODE 0: y = f(x) for the tile.  MODE 1: gradient of the loss over the tile
// (scaled by 1/n_total) into part[blockIdx.x][NP] and the tile's summed
// squared error into sse[blockIdx.x].  `idx` (optional) maps tile rows to
// sample rows.
template <int MODE>
__global__ void __launch_bounds__(256) k_mlp_tile(const float *__restrict__ P, const float *__restrict__ X,
                                                  const float *__restrict__ T, const int32_t *__restrict__ idx,
                                                  int n, int row0, float inv_n, float *__restrict__ Y,
                                                  float *__restrict__ part, float *__restrict__ sse) {
  __shared__ float L[L_END];
  const int tid = threadIdx.x;
  const int base = row0 + blockIdx.x * TILE;
  const int nvalid = max(0, min(TILE, n - base));
  // weights into LDS (padded rows)
  for (int k = tid; k < NH * NI; k += 256) L[L_W1 + (k / NI) * S1 + k % NI] = P[OW1 + k];
  for (int k = tid; k < NH * NH; k += 256) L[L_W2 + (k / NH) * S2 + k % NH] = P[OW2 + k];
  for (int k = tid; k < NO * NH; k += 256) L[L_W3 + (k / NH) * S2 + k % NH] = P[OW3 + k];
  if (tid < NH) {
    L[L_B1 + tid] = P[OB1 + tid];
    L[L_B2 + tid] = P[OB2 + tid];
  }
  if (tid < NO) L[L_B3 + tid] = P[OB3 + tid];
  for (int k = tid; k < TILE * NI; k += 256) {
    const int s = k / NI;
    float v = 0.f;
    if (s < nvalid) {
      const int row = idx ? idx[base + s] : base + s;
      v = X[(size_t)row * NI + k % NI];
    }
    L[L_X + k] = v;
  }
  __syncthreads();
  // hidden layer 1: thread -> (sample s, unit i), consecutive i across lanes
  for (int o = tid; o < TILE * NH; o += 256) {
    const int s = o / NH, i = o % NH;
    float a = L[L_B1 + i];
#pragma unroll 8
    for (int j = 0; j < NI; ++j) a = fmaf(L[L_W1 + i * S1 + j], L[L_X + s * NI + j], a);
    L[L_H1 + o] = fmaxf(a, 0.f);
  }
  __syncthreads();
  for (int o = tid; o < TILE * NH; o += 256) {
    const int s = o / NH, k = o % NH;
    float a = L[L_B2 + k];
#pragma unroll 8
    for (int j = 0; j < NH; ++j) a = fmaf(L[L_W2 + k * S2 + j], L[L_H1 + s * NH + j], a);
    L[L_H2 + o] = fmaxf(a, 0.f);
  }
  __syncthreads();
  if (tid < TILE * NO) {
    const int s = tid / NO, q = tid % NO;
    float a = L[L_B3 + q];
#pragma unroll 8
    for (int k = 0; k < NH; ++k) a = fmaf(L[L_W3 + q * S2 + k], L[L_H2 + s * NH + k], a);
    L[L_Y + tid] = a;
    if (MODE == 0 && s < nvalid) {
      const int row = idx ? idx[base + s] : base + s;
      Y[(size_t)row * NO + q] = a;
    }
    if (MODE == 1) {
      float d = 0.f;
      if (s < nvalid) {
        const int row = idx ? idx[base + s] : base + s;
        d = a - T[(size_t)row * NO + q];
      }
      L[L_DY + tid] = 2.f * d * inv_n;
      L[L_D1 + tid] = d * d; // squared error, summed below before the D1 area is used
    }
  }
  if (MODE == 0) return;
  __syncthreads();
  if (tid == 0) {
    float e = 0.f;
    for (int k = 0; k < TILE * NO; ++k) e += L[L_D1 + k];
    sse[blockIdx.x] = e;
  }
  __syncthreads();
  float *G = part + (size_t)blockIdx.x * NP;
  // output layer: dW3[q][k] = sum_s dY[s][q] H2[s][k]; dH2 -> D2 with the ReLU mask
  for (int o = tid; o < NO * NH; o += 256) {
    const int q = o / NH, k = o % NH;
    float g = 0.f;
    for (int s = 0; s < TILE; ++s) g = fmaf(L[L_DY + s * NO + q], L[L_H2 + s * NH + k], g);
    G[OW3 + o] = g;
  }
  if (tid < NO) {
    float g = 0.f;
    for (int s = 0; s < TILE; ++s) g += L[L_DY + s * NO + tid];
    G[OB3 + tid] = g;
  }
  for (int o = tid; o < TILE * NH; o += 256) {
    const int s = o / NH, k = o % NH;
    float g = 0.f;
#pragma unroll
    for (int q = 0; q < NO; ++q) g = fmaf(L[L_W3 + q * S2 + k], L[L_DY + s * NO + q], g);
    L[L_D2 + o] = L[L_H2 + o] > 0.f ? g : 0.f;
  }
  __syncthreads();
  // layer 2: dW2[k][j] = sum_s D2[s][k] H1[s][j]; dH1[s][j] = sum_k W2[k][j] D2[s][k], masked
  for (int o = tid; o < NH * NH; o += 256) {
    const int k = o / NH, j = o % NH;
    float g = 0.f;
    for (int s = 0; s < TILE; ++s) g = fmaf(L[L_D2 + s * NH + k], L[L_H1 + s * NH + j], g);
    G[OW2 + o] = g;
  }
  if (tid < NH) {
    float g = 0.f;
    for (int s = 0; s < TILE; ++s) g += L[L_D2 + s * NH + tid];
    G[OB2 + tid] = g;
  }
  for (int o = tid; o < TILE * NH; o += 256) {
    const int s = o / NH, j = o % NH;
    float g = 0.f;
#pragma unroll 8
    for (int k = 0; k < NH; ++k) g = fmaf(L[L_W2 + k * S2 + j], L[L_D2 + s * NH + k], g);
    L[L_D1 + o] = L[L_H1 + o] > 0.f ? g : 0.f;
  }
  __syncthreads();
  // layer 1: dW1[i][j] = sum_s D1[s][i] X[s][j]
  for (int o = tid; o < NH * NI; o += 256) {
    const int i = o / NI, j = o % NI;
    float g = 0.f;
    for (int s = 0; s < TILE; ++s) g = fmaf(L[L_D1 + s * NH + i], L[L_X + s * NI + j], g);
    G[OW1 + o] = g;
  }
  if (tid < NH) {
    float g = 0.f;
    for (int s = 0; s < TILE; ++s) g += L[L_D1 + s * NH + tid];
    G[OB1 + tid] = g;
  }
}

// grad[p] = sum over tiles (in tile order) of part[tile][p]
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *__restrict__ part, int ntiles, float *__restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NP) return;
  float g = 0.f;
  for (int b = 0; b < ntiles; ++b) g += part[(size_t)b * NP + p];
  grad[p] = g;
}

// Adam (Kingma & Ba 2015), step t >= 1, bias-corrected.
__global__ void __launch_bounds__(256) k_mlp_adam(float *__restrict__ P, const float *__restrict__ G,
                                                  float *__restrict__ M, float *__restrict__ V, float lr, float b1,
                                                  float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NP) return;
  const float g = G[p];
  const float m = b1 * M[p] + (1.f - b1) * g;
  const float v = b2 * V[p] + (1.f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

extern "C" {

int e13k_mlp_forward(const float *P, const float *X, const int32_t *idx, int n, float *Y, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_mlp_tile<0>, dim3((n + TILE - 1) / TILE), dim3(256), 0, s, P, X, (const float *)nullptr, idx,
                     n, 0, 0.f, Y, (float *)nullptr, (float *)nullptr);
  return (int)hipGetLastError();
}

// Gradient of the loss over rows [row0, row0 + m) of the (optionally
// index-mapped) sample set, scaled by inv_n; part needs ceil(m/32) x 6402
// floats and sse ceil(m/32) floats.
int e13k_mlp_grad(const float *P, const float *X, const float *T, const int32_t *idx, int row0, int m, float inv_n,
                  float *part, float *sse, float *grad, hipStream_t s) {
  if (m <= 0) return 0;
  const int ntiles = (m + TILE - 1) / TILE;
  hipLaunchKernelGGL(k_mlp_tile<1>, dim3(ntiles), dim3(256), 0, s, P, X, T, idx, row0 + m, row0, inv_n,
                     (float *)nullptr, part, sse);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((NP + 255) / 256), dim3(256), 0, s, (const float *)part, ntiles, grad);
  return (int)hipGetLastError();
}

int e13k_mlp_adam(float *P, const float *G, float *M, float *V, float lr, float b1, float b2, float eps, float c1,
                  float c2, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_adam, dim3((NP + 255) / 256), dim3(256), 0, s, P, G, M, V, lr, b1, b2, eps, c1, c2);
  return (int)hipGetLastError();
}

} // extern "C"
