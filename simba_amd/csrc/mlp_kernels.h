// mlp_kernels.h -- launch interface of the block-proposal MLP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mba {

constexpr int kMlpIn = 256;
// features: kFeatHashed hashed sample-bit buckets, then a one-hot of the size
// whose blocks are being ordered (slot kFeatHashed + min(size, 16) - 1)
constexpr int kFeatHashed = 240;
constexpr int kFeatSizes = kMlpIn - kFeatHashed;
// bucket of a hashed pattern: multiply-shift range reduction to kFeatHashed
__host__ __device__ inline uint32_t feat_bucket(uint32_t h) {
  return (uint32_t)(((uint64_t)h * (uint64_t)kFeatHashed) >> 32);
}
__host__ __device__ inline int feat_size_slot(int size) {
  return kFeatHashed + (size < 1 ? 0 : size > kFeatSizes ? kFeatSizes - 1 : size - 1);
}
constexpr int kMlpHidden = 512;
constexpr int kMlpOut = 128;   // class = op * 16 + min(left size, 15)

// C[M x N] (row stride ldc) = A . B with A(i, k) = A[i*sa_i + k*sa_k] and
// B(k, j) = B[k*sb_k + j*sb_j]; epilogue: + bias[j], ReLU, or zero where
// gate[i][j] <= 0 (ReLU backward).  M, N, K multiples of 16.
struct GemmArgs {
  const float *A, *B;
  float *C;
  int M, N, K;
  int64_t sa_i, sa_k, sb_k, sb_j, ldc;
  const float *bias;
  const float *gate;
  int relu;
};

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s);
hipError_t launch_softmax_ce(const float *Z, const int32_t *labels, int B, int K, float inv_n,
                             float *dZ, float *loss_rows, hipStream_t s);
hipError_t launch_colsum(const float *D, int rows, int cols, float *out, hipStream_t s);
// One instance of a feature batch: its [sample][var] inputs and outputs in
// the packed arrays.
struct FeatInst {
  uint64_t in_off, out_off;
  uint32_t n_samples, n_vars, width, pad;
};
// X[i * 256 + b] = the 256 hashed features of instance i (the host
// instance_features, bit for bit), instances [0, count).
hipError_t launch_features(const uint64_t *d_inputs, const uint64_t *d_outputs,
                           const FeatInst *d_fi, uint32_t count, float *d_x, hipStream_t s);
hipError_t launch_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t s);

}  // namespace mba
