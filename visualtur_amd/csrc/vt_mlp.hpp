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

// offsets (in floats) of the biases and the logit weights in the fp32 parameters
struct ScoreParams {
    int b1, b2, b3, w4, b4;
};
// optional filtering of the scored candidates against the page table
// (table == nullptr: logits only); codes[i] - base indexes the table
struct ScoreFilter {
    const int32_t* table = nullptr;
    uint64_t base = 0;
    float zmin = 0.0f;
    float* slot_score = nullptr;   // logit of each resident candidate, by slot
    uint32_t* nwant = nullptr;     // absent candidates with logit >= zmin: count,
    int32_t* want_idx = nullptr;   // candidate index
    float* want_z = nullptr;       // and logit, in no particular order
};
// WT = the hidden layers' BF16 weights transposed to [n][k] (256x32, 256x256, 256x256)
hipError_t launch_pack_wt(const bf16* W1, const bf16* W2, const bf16* W3, bf16* WT, hipStream_t s);
// fused features + MLP forward of N candidate bricks: logits into z[N]
hipError_t launch_score(const Camera& cam, int n, int level_c, const uint64_t* codes, int N, const bf16* WT,
                        const float* P, const ScoreParams& sp, float* z, const ScoreFilter& f, hipStream_t s);
hipError_t launch_fill_f32(float* x, int n, float v, hipStream_t s);

}  // namespace vt
