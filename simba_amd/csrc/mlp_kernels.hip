// mlp_kernels.hip -- the learned block-proposal network on gfx950: a 2-layer
// MLP (256 -> 512 -> 128), BF16 MFMA operands with FP32 accumulation, and
// its training kernels (softmax cross-entropy, backward GEMMs, Adam).
//
// GEMM: C[M x N] = A[M x K] . B[K x N] with arbitrary element strides, so
// the transposed products of the backward pass need no copies.  Operands are
// held in FP32 and rounded to BF16 (round to nearest even) as they are
// staged into the MFMA registers; accumulation is FP32.  One wave computes
// one 16 x 16 tile with v_mfma_f32_16x16x16_bf16: per 16-deep K step lane l
// supplies A[l % 16][k0 + 4(l/16) .. +3] and B[k0 + 4(l/16) .. +3][l % 16],
// and receives D[4(l/16) + i][l % 16] in its i-th accumulator.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp_kernels.h"

namespace mba {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);                                // nearest even
  return (short)(u >> 16);
}

__global__ void __launch_bounds__(256)
k_gemm_bf16(GemmArgs g) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tiles_n = g.N / 16;
  const int tile = blockIdx.x * 4 + wave;
  if (tile >= (g.M / 16) * tiles_n) return;
  const int ti = tile / tiles_n, tj = tile % tiles_n;
  const int r = lane & 15, q = lane >> 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    bf16x4 a, b;
    for (int e = 0; e < 4; ++e) {
      const int k = k0 + 4 * q + e;
      a[e] = f32_to_bf16(g.A[(int64_t)(ti * 16 + r) * g.sa_i + (int64_t)k * g.sa_k]);
      b[e] = f32_to_bf16(g.B[(int64_t)k * g.sb_k + (int64_t)(tj * 16 + r) * g.sb_j]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  const int col = tj * 16 + r;
  for (int i = 0; i < 4; ++i) {
    const int row = ti * 16 + 4 * q + i;
    float v = acc[i];
    if (g.bias) v += g.bias[col];
    if (g.relu) v = v > 0.f ? v : 0.f;
    if (g.gate) v = g.gate[(int64_t)row * g.ldc + col] > 0.f ? v : 0.f;
    g.C[(int64_t)row * g.ldc + col] = v;
  }
}

// Softmax cross-entropy over rows of Z [B x K]: loss (mean over the rows
// with label >= 0) and dZ = (softmax - onehot) / n_labelled.
__global__ void __launch_bounds__(64)
k_softmax_ce(const float *Z, const int32_t *labels, int B, int K, float inv_n, float *dZ,
             float *loss_rows) {
  const int row = blockIdx.x;
  if (row >= B) return;
  const float *z = Z + (int64_t)row * K;
  float m = -3.0e38f;
  for (int k = threadIdx.x; k < K; k += 64) m = fmaxf(m, z[k]);
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  float s = 0.f;
  for (int k = threadIdx.x; k < K; k += 64) s += __expf(z[k] - m);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  const int y = labels[row];
  const float lse = m + __logf(s);
  for (int k = threadIdx.x; k < K; k += 64) {
    const float p = __expf(z[k] - lse);
    dZ[(int64_t)row * K + k] = y >= 0 ? (p - (k == y ? 1.f : 0.f)) * inv_n : 0.f;
  }
  if (threadIdx.x == 0) loss_rows[row] = y >= 0 ? (lse - z[y]) : 0.f;
}

// db[j] = sum_i D[i][j]
__global__ void k_colsum(const float *D, int rows, int cols, float *out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  float s = 0.f;
  for (int i = 0; i < rows; ++i) s += D[(int64_t)i * cols + j];
  out[j] = s;
}

// Adam (Kingma & Ba, ICLR 2015) on FP32 master weights.
__global__ void k_adam(float *w, const float *g, float *m, float *v, int n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s) {
  if (g.M % 16 || g.N % 16 || g.K % 16) return hipErrorInvalidValue;
  const int tiles = (g.M / 16) * (g.N / 16);
  hipLaunchKernelGGL(k_gemm_bf16, dim3((tiles + 3) / 4), dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t launch_softmax_ce(const float *Z, const int32_t *labels, int B, int K, float inv_n,
                             float *dZ, float *loss_rows, hipStream_t s) {
  hipLaunchKernelGGL(k_softmax_ce, dim3(B), dim3(64), 0, s, Z, labels, B, K, inv_n, dZ,
                     loss_rows);
  return hipGetLastError();
}

hipError_t launch_colsum(const float *D, int rows, int cols, float *out, hipStream_t s) {
  hipLaunchKernelGGL(k_colsum, dim3((cols + 255) / 256), dim3(256), 0, s, D, rows, cols, out);
  return hipGetLastError();
}

// Features of a batch of instances, one workgroup (one wave) per instance,
// lane k = bit position k: the pattern (output bit, bits of the first five
// variables) of every sample at bit k, hashed with the bit group and the
// variable count into one of kFeatHashed buckets (a histogram in LDS),
// scaled by 8 / (samples * width); the size slots stay zero.  Counts stay
// exact in float, so the result equals the host's running float sums.
__device__ __forceinline__ uint32_t feat_mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x85ebca6bu;
  x ^= x >> 13;
  x *= 0xc2b2ae35u;
  x ^= x >> 16;
  return x;
}

__global__ void __launch_bounds__(64)
k_features(const uint64_t *__restrict__ inputs, const uint64_t *__restrict__ outputs,
           const FeatInst *__restrict__ fi, uint32_t count, float *__restrict__ x) {
  __shared__ uint32_t hist[kMlpIn];
  const uint32_t i = blockIdx.x;
  if (i >= count) return;
  const FeatInst f = fi[i];
  for (int b = threadIdx.x; b < kMlpIn; b += 64) hist[b] = 0u;
  __syncthreads();
  const uint32_t k = threadIdx.x;
  if (k < f.width) {
    const uint32_t g = k < 8 ? k : k < 16 ? 8u : k < 32 ? 9u : 10u;
    const uint32_t nv5 = f.n_vars < 5 ? f.n_vars : 5u;
    for (uint32_t smp = 0; smp < f.n_samples; ++smp) {
      uint32_t p = (uint32_t)((outputs[f.out_off + smp] >> k) & 1u);
      const uint64_t *in = inputs + f.in_off + (size_t)smp * f.n_vars;
      for (uint32_t v = 0; v < nv5; ++v) p |= (uint32_t)((in[v] >> k) & 1u) << (v + 1);
      atomicAdd(&hist[feat_bucket(feat_mix(p * 64u + g + 1024u * f.n_vars))], 1u);
    }
  }
  __syncthreads();
  const float scale = 8.f / (float)(f.n_samples * f.width);
  for (int b = threadIdx.x; b < kMlpIn; b += 64) x[(size_t)i * kMlpIn + b] = (float)hist[b] * scale;
}

hipError_t launch_features(const uint64_t *d_inputs, const uint64_t *d_outputs,
                           const FeatInst *d_fi, uint32_t count, float *d_x, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_features, dim3(count), dim3(64), 0, s, d_inputs, d_outputs, d_fi, count,
                     d_x);
  return hipGetLastError();
}

hipError_t launch_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2,
                     eps, c1, c2);
  return hipGetLastError();
}

}  // namespace mba
