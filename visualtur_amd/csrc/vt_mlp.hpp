// vt_mlp.hpp -- BF16 MFMA MLP kernels (brick-visibility prefetcher).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vt_device.hpp"

namespace vt {

typedef __bf16 bf16;

enum GemmEpi : int { EPI_F32 = 0, EPI_ATOMIC_F32 = 1, EPI_BIAS_RELU_BF16 = 2, EPI_MASK_BF16 = 3 };

// C = op(A) op(B); A logical MxK (ta: stored KxM), B logical KxN (tb: stored NxK)
hipError_t launch_gemm(bool ta, bool tb, int epi, const bf16* A, int lda, const bf16* B, int ldb, float* Cf,
                       bf16* Cb, int ldc, int M, int N, int K, int splitk, const float* bias, const bf16* mask,
                       int ldm, hipStream_t s);
hipError_t launch_logit(const bf16* H, const float* w, const float* b, int N, int width, float* z, hipStream_t s);
hipError_t launch_bce(const float* z, const float* y, int N, float pos_weight, float* dz, float* loss_sum,
                      hipStream_t s);
hipError_t launch_outer_mask(const float* dz, const float* w, const bf16* H, int N, int width, bf16* dH,
                             hipStream_t s);
hipError_t launch_colsum(const bf16* X, const float* sc, int N, int width, float* out, hipStream_t s);
hipError_t launch_sum(const float* x, int N, float* out, hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, bf16* pb, int n, float lr, float b1, float b2,
                       float eps, int step, hipStream_t s);
hipError_t launch_to_bf16(const float* x, bf16* y, size_t n, hipStream_t s);
hipError_t launch_sigmoid(float* z, int n, hipStream_t s);
// X[N x 32] bf16 (features 0..11, zero padding) from fp32 x[N x 12]
hipError_t launch_pack_x(const float* x, int N, bf16* X, hipStream_t s);
// brick features for a camera: x[N x 12] fp32
hipError_t launch_features(const Camera& cam, int n, int level_c, const uint64_t* codes, int N, float* x,
                           hipStream_t s);

}  // namespace vt
