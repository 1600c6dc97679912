// vt_mlp.hip -- BF16 MFMA kernels of the brick-visibility MLP (no reference
// counterpart).  Network: 12 features -> 256 -> 256 -> 256 -> 1 logit, ReLU
// hidden layers; BF16 weights and activations, FP32 accumulation, FP32
// master weights updated by Adam.
//
// GEMM: C[MxN] = op(A)[MxK] * op(B)[KxN] with v_mfma_f32_16x16x32_bf16.
// A 256-thread workgroup owns a 64x64 tile of C; each wave64 owns 32x32 of
// it as 2x2 MFMA blocks.  K advances 32 at a time through LDS.  MFMA operand
// layout (16x16x32): lane l supplies A[l%16][8*(l/16) .. +7] and
// B[8*(l/16) .. +7][l%16]; it receives D[4*(l/16) .. +3][l%16].
#include "vt_mlp.hpp"

namespace vt {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 32, LDSK = TK + 8;   // +8: 80-byte rows, 16-byte aligned

// eight consecutive elements of an operand row, `valid` of them in range
// (0..8): one 16-byte load when all are in range and `vec` (16-byte aligned
// rows), else element by element with zeros past the end
__device__ __forceinline__ bf16x8 load8(const bf16* __restrict__ p, int valid, bool vec)
{
    if (vec && valid == 8) return *reinterpret_cast<const bf16x8*>(p);
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = i < valid ? p[i] : (bf16)0.0f;
    return v;
}

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B,
                                              int ldb, float* __restrict__ Cf, bf16* __restrict__ Cb, int ldc,
                                              int M, int N, int K, int kchunk, const float* __restrict__ bias,
                                              const bf16* __restrict__ mask, int ldm, bool vec)
{
    __shared__ __attribute__((aligned(16))) bf16 As[TM][LDSK];   // As[m][k]
    __shared__ __attribute__((aligned(16))) bf16 Bs[TN][LDSK];   // Bs[n][k] (B transposed)
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        // stage the 64x32 A and B tiles: each thread moves one run of eight
        // elements along the operand's contiguous dimension (k for a
        // row-major operand, m or n for a transposed one)
        if (!TA) {
            const int r = tid >> 2, c = (tid & 3) * 8, gm = m0 + r, gk = k0 + c;
            const int valid = gm < M ? max(0, min(8, kend - gk)) : 0;
            *reinterpret_cast<bf16x8*>(&As[r][c]) = load8(A + (size_t)gm * lda + gk, valid, vec);
        } else {
            // a wave's lanes take 32 k rows x 2 runs, so each of the eight
            // column stores below spreads over 32 LDS banks
            const int kr = tid & 31, c = (tid >> 5) * 8, gk = k0 + kr, gm = m0 + c;
            const int valid = gk < kend ? max(0, min(8, M - gm)) : 0;
            const bf16x8 v = load8(A + (size_t)gk * lda + gm, valid, vec);
#pragma unroll
            for (int i = 0; i < 8; ++i) As[c + i][kr] = v[i];
        }
        if (TB) {
            const int r = tid >> 2, c = (tid & 3) * 8, gn = n0 + r, gk = k0 + c;
            const int valid = gn < N ? max(0, min(8, kend - gk)) : 0;
            *reinterpret_cast<bf16x8*>(&Bs[r][c]) = load8(B + (size_t)gn * ldb + gk, valid, vec);
        } else {
            const int kr = tid & 31, c = (tid >> 5) * 8, gk = k0 + kr, gn = n0 + c;
            const int valid = gk < kend ? max(0, min(8, N - gn)) : 0;
            const bf16x8 v = load8(B + (size_t)gk * ldb + gn, valid, vec);
#pragma unroll
            for (int i = 0; i < 8; ++i) Bs[c + i][kr] = v[i];
        }
        __syncthreads();
        const int kk = 8 * (lane >> 4);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm + 16 * i + (lane & 15)][kk]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn + 16 * j + (lane & 15)][kk]);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int row = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
                int col = n0 + wn + 16 * j + (lane & 15);
                if (row >= M || col >= N) continue;
                float x = acc[i][j][v];
                size_t o = (size_t)row * ldc + col;
                if (EPI == EPI_F32) {
                    Cf[o] = x;
                } else if (EPI == EPI_ATOMIC_F32) {
                    atomicAdd(&Cf[o], x);
                } else if (EPI == EPI_BIAS_RELU_BF16) {
                    x = x + bias[col];
                    Cb[o] = (bf16)(x > 0.0f ? x : 0.0f);
                } else if (EPI == EPI_MASK_BF16) {
                    float mk = (float)mask[(size_t)row * ldm + col];
                    Cb[o] = (bf16)(mk > 0.0f ? x : 0.0f);
                }
            }
}

// logit z[i] = H[i,:] . w + b, one wave per row
__global__ __launch_bounds__(256) void k_logit(const bf16* __restrict__ H, const float* __restrict__ w,
                                               const float* __restrict__ b, int N, int width, float* __restrict__ z)
{
    int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (row >= N) return;
    float s = 0.0f;
    for (int j = lane; j < width; j += 64) s += (float)H[(size_t)row * width + j] * w[j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) z[row] = s + b[0];
}

// weighted BCE-with-logits: loss_i = w_i (softplus(z) - y z), dz_i = w_i (sigmoid(z) - y) / N
__global__ void k_bce(const float* __restrict__ z, const float* __restrict__ y, int N, float pos_weight,
                      float* __restrict__ dz, float* __restrict__ loss_sum)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < N) {
        float zi = z[i], yi = y[i];
        float wi = yi > 0.5f ? pos_weight : 1.0f;
        float sp = zi > 0.0f ? zi + log1pf(expf(-zi)) : log1pf(expf(zi));
        l = wi * (sp - yi * zi);
        float sg = 1.0f / (1.0f + expf(-zi));
        dz[i] = wi * (sg - yi) / (float)N;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss_sum, l);
}

// dH[i][j] = dz[i] * w[j] where H[i][j] > 0 (ReLU derivative), as BF16
__global__ void k_outer_mask(const float* __restrict__ dz, const float* __restrict__ w, const bf16* __restrict__ H,
                             int N, int width, bf16* __restrict__ dH)
{
    size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)N * width) return;
    int i = (int)(e / width), j = (int)(e % width);
    float h = (float)H[e];
    dH[e] = (bf16)(h > 0.0f ? dz[i] * w[j] : 0.0f);
}

// column sums out[j] (+)= sum_i X[i][j] * (s ? s[i] : 1): one column per
// thread, kColsumRows rows per workgroup (enough workgroups to cover the
// chip at a training batch of 8192), four independent partial sums per
// thread so that four loads are in flight
constexpr int kColsumRows = 32;

__global__ void k_colsum(const bf16* __restrict__ X, const float* __restrict__ s, int N, int width,
                         float* __restrict__ out)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int r0 = blockIdx.y * kColsumRows, r1 = min(N, r0 + kColsumRows);
    if (j >= width) return;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    int i = r0;
    for (; i + 4 <= r1; i += 4) {
        a0 += (float)X[(size_t)i * width + j] * (s ? s[i] : 1.0f);
        a1 += (float)X[(size_t)(i + 1) * width + j] * (s ? s[i + 1] : 1.0f);
        a2 += (float)X[(size_t)(i + 2) * width + j] * (s ? s[i + 2] : 1.0f);
        a3 += (float)X[(size_t)(i + 3) * width + j] * (s ? s[i + 3] : 1.0f);
    }
    for (; i < r1; ++i) a0 += (float)X[(size_t)i * width + j] * (s ? s[i] : 1.0f);
    atomicAdd(&out[j], (a0 + a1) + (a2 + a3));
}

__global__ void k_sum(const float* __restrict__ x, int N, float* __restrict__ out)
{
    float a = 0.0f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) a += x[i];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, a);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, bf16* __restrict__ pb, int n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (pb) pb[i] = (bf16)pi;
}

__global__ void k_to_bf16(const float* __restrict__ x, bf16* __restrict__ y, size_t n)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = (bf16)x[i];
}

__global__ void k_sigmoid(float* __restrict__ z, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = 1.0f / (1.0f + expf(-z[i]));
}

__global__ void k_pack_x(const float* __restrict__ x, int N, bf16* __restrict__ X)
{
    size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= (size_t)N * 32) return;
    int i = (int)(e / 32), j = (int)(e % 32);
    X[e] = (bf16)(j < 12 ? x[(size_t)i * 12 + j] : 0.0f);
}

// 12 features of (camera, brick): camera position and brick centre in
// units of the cube (centred), view direction, brick level, and the brick
// centre's image-plane position (u, v)
__device__ __forceinline__ void brick_features(const Camera& cam, int n, int level_c, uint64_t code, float* o)
{
    const float D = (float)(1 << n);
    int bx, by, bz;
    node_min(code, level_c, n, bx, by, bz);
    const float half = 0.5f * (float)(1 << (n - level_c));
    float c0 = (float)bx + half, c1 = (float)by + half, c2 = (float)bz + half;
    float d0 = c0 - cam.pos[0], d1 = c1 - cam.pos[1], d2 = c2 - cam.pos[2];
    // image-plane position of the brick centre: (u, v) in [-1, 1] on screen
    // (the ray of pixel (i, j) has u = 2(j+0.5)/W - 1, v = 2(i+0.5)/H - 1),
    // clamped to [-4, 4]; 4 for a centre behind the camera
    float dl = d0 * cam.look[0] + d1 * cam.look[1] + d2 * cam.look[2];
    float dr = d0 * cam.right[0] + d1 * cam.right[1] + d2 * cam.right[2];
    float du = d0 * cam.up[0] + d1 * cam.up[1] + d2 * cam.up[2];
    float u = 4.0f, v = 4.0f;
    if (dl > 0.0f) {
        u = fminf(fmaxf(dr / dl * (2.0f * cam.distance / cam.w), -4.0f), 4.0f);
        v = fminf(fmaxf(du / dl * (2.0f * cam.distance / cam.h), -4.0f), 4.0f);
    }
    o[0] = cam.pos[0] / D - 0.5f;
    o[1] = cam.pos[1] / D - 0.5f;
    o[2] = cam.pos[2] / D - 0.5f;
    o[3] = cam.look[0];
    o[4] = cam.look[1];
    o[5] = cam.look[2];
    o[6] = c0 / D - 0.5f;
    o[7] = c1 / D - 0.5f;
    o[8] = c2 / D - 0.5f;
    o[9] = (float)level_c / 10.0f;
    o[10] = u;
    o[11] = v;
}

__global__ void k_features(Camera cam, int n, int level_c, const uint64_t* __restrict__ codes, int N,
                           float* __restrict__ x)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    brick_features(cam, n, level_c, codes[i], x + (size_t)i * 12);
}

// WT[n][k] = W[k][n] for the three hidden layers' BF16 weights (W1 32x256,
// W2 and W3 256x256, row-major [k][n] in Pb), packed back to back
__global__ void k_pack_wt(const bf16* __restrict__ W1, const bf16* __restrict__ W2, const bf16* __restrict__ W3,
                          bf16* __restrict__ WT)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int n1 = 256 * 32, n2 = 256 * 256;
    if (e < n1) {
        int nn = e / 32, k = e % 32;
        WT[e] = W1[k * 256 + nn];
    } else if (e < n1 + 2 * n2) {
        int f = (e - n1) % n2, nn = f / 256, k = f % 256;
        WT[e] = (e < n1 + n2 ? W2 : W3)[k * 256 + nn];
    }
}

// Scoring: features -> three 256-wide layers -> logit for 64 candidates per
// 256-thread workgroup.  The activations stay in LDS between layers (one
// 64x256 BF16 buffer, overwritten in place once every wave holds its layer's
// accumulators in registers); the transposed weights are read as MFMA B operands
// straight from L2 (16 bytes per lane).  Every output is accumulated with
// v_mfma_f32_16x16x32_bf16 over k = 0, 32, 64, ... as in k_gemm, and the bias,
// ReLU, BF16 rounding and logit reduction are k_gemm's and k_logit's, so the
// logits equal the layer-by-layer path's bit for bit.
constexpr int SR = 64, SH = 256, LDH = SH + 8;

// H[r][c] = bf16(relu(sum_k H[r][k] WT[c][k] + bias[c])) for the 64 columns
// of this wave; the first barrier keeps every wave's reads of the layer's
// input ahead of the writes of its output
template <int K>
__device__ __forceinline__ void score_layer(bf16 (*H)[LDH], const bf16* __restrict__ WT,
                                            const float* __restrict__ bias, int wave, int lane)
{
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int kq = 8 * (lane >> 4), l16 = lane & 15, c0 = 64 * wave;
#pragma unroll 2
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 b[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            b[j] = *reinterpret_cast<const bf16x8*>(WT + (size_t)(c0 + 16 * j + l16) * K + k0 + kq);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            bf16x8 a = *reinterpret_cast<const bf16x8*>(&H[16 * i + l16][k0 + kq]);
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = c0 + 16 * j + l16;
            const float bc = bias[col];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                float x = acc[i][j][v] + bc;
                H[16 * i + 4 * (lane >> 4) + v][col] = (bf16)(x > 0.0f ? x : 0.0f);
            }
        }
}

// Writes z[i] for the N candidates.  With a page table: a resident
// candidate's logit goes to slot_score[its slot]; an absent one at or above
// zmin is appended to (want_idx, want_z) through the counter *nwant.
__global__ __launch_bounds__(256) void k_score(Camera cam, int n, int level_c, const uint64_t* __restrict__ codes,
                                               int N, const bf16* __restrict__ WT, const float* __restrict__ P,
                                               ScoreParams sp, float* __restrict__ z, ScoreFilter f)
{
    __shared__ __attribute__((aligned(16))) bf16 H[SR][LDH];
    const int r0 = blockIdx.x * SR;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid < SR) {
        float x[12];
        const int i = r0 + tid;
        if (i < N) {
            brick_features(cam, n, level_c, codes[i], x);
        } else {
#pragma unroll
            for (int j = 0; j < 12; ++j) x[j] = 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 32; ++j) H[tid][j] = (bf16)(j < 12 ? x[j] : 0.0f);
    }
    __syncthreads();
    score_layer<32>(H, WT, P + sp.b1, wave, lane);
    __syncthreads();
    score_layer<SH>(H, WT + SH * 32, P + sp.b2, wave, lane);
    __syncthreads();
    score_layer<SH>(H, WT + SH * 32 + SH * SH, P + sp.b3, wave, lane);
    __syncthreads();
    const float* w4 = P + sp.w4;
    for (int r = 16 * wave; r < 16 * wave + 16; ++r) {
        float s = 0.0f;
        for (int j = lane; j < SH; j += 64) s += (float)H[r][j] * w4[j];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        const int i = r0 + r;
        if (lane != 0 || i >= N) continue;
        const float zi = s + P[sp.b4];
        z[i] = zi;
        if (!f.table) continue;
        const int32_t slot = f.table[codes[i] - f.base];
        if (slot >= 0) {
            f.slot_score[slot] = zi;
        } else if (zi >= f.zmin) {
            const uint32_t k = atomicAdd(f.nwant, 1u);
            f.want_idx[k] = i;
            f.want_z[k] = zi;
        }
    }
}

__global__ void k_fill_f32(float* __restrict__ x, int n, float v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

}  // namespace

hipError_t launch_pack_x(const float* x, int N, bf16* X, hipStream_t s)
{
    size_t total = (size_t)N * 32;
    hipLaunchKernelGGL(k_pack_x, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, x, N, X);
    return hipGetLastError();
}

hipError_t launch_pack_wt(const bf16* W1, const bf16* W2, const bf16* W3, bf16* WT, hipStream_t s)
{
    constexpr int total = 256 * 32 + 2 * 256 * 256;
    hipLaunchKernelGGL(k_pack_wt, dim3((total + 255) / 256), dim3(256), 0, s, W1, W2, W3, WT);
    return hipGetLastError();
}

hipError_t launch_score(const Camera& cam, int n, int level_c, const uint64_t* codes, int N, const bf16* WT,
                        const float* P, const ScoreParams& sp, float* z, const ScoreFilter& f, hipStream_t s)
{
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_score, dim3((N + SR - 1) / SR), dim3(256), 0, s, cam, n, level_c, codes, N, WT, P, sp, z, f);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float* x, int n, float v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f32, dim3((n + 255) / 256), dim3(256), 0, s, x, n, v);
    return hipGetLastError();
}

hipError_t launch_features(const Camera& cam, int n, int level_c, const uint64_t* codes, int N, float* x,
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_features, dim3((N + 255) / 256), dim3(256), 0, s, cam, n, level_c, codes, N, x);
    return hipGetLastError();
}

template <bool TA, bool TB, int EPI>
static hipError_t gemm_t(const bf16* A, int lda, const bf16* B, int ldb, float* Cf, bf16* Cb, int ldc, int M, int N,
                         int K, int splitk, const float* bias, const bf16* mask, int ldm, hipStream_t s)
{
    int kchunk = (K + splitk - 1) / splitk;
    kchunk = (kchunk + TK - 1) / TK * TK;
    int nz = (K + kchunk - 1) / kchunk;
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, nz);
    // 16-byte runs need rows that start on 16 bytes
    const bool vec = lda % 8 == 0 && ldb % 8 == 0 && ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0;
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, s, A, lda, B, ldb, Cf, Cb, ldc, M, N, K, kchunk,
                       bias, mask, ldm, vec);
    return hipGetLastError();
}

hipError_t launch_gemm(bool ta, bool tb, int epi, const bf16* A, int lda, const bf16* B, int ldb, float* Cf,
                       bf16* Cb, int ldc, int M, int N, int K, int splitk, const float* bias, const bf16* mask,
                       int ldm, hipStream_t s)
{
#define VT_GEMM_CASE(TA_, TB_, E_) \
    if (ta == TA_ && tb == TB_ && epi == E_) \
        return gemm_t<TA_, TB_, E_>(A, lda, B, ldb, Cf, Cb, ldc, M, N, K, splitk, bias, mask, ldm, s);
    VT_GEMM_CASE(false, false, EPI_F32)
    VT_GEMM_CASE(false, false, EPI_BIAS_RELU_BF16)
    VT_GEMM_CASE(false, true, EPI_F32)
    VT_GEMM_CASE(false, true, EPI_MASK_BF16)
    VT_GEMM_CASE(true, false, EPI_F32)
    VT_GEMM_CASE(true, false, EPI_ATOMIC_F32)
#undef VT_GEMM_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_logit(const bf16* H, const float* w, const float* b, int N, int width, float* z, hipStream_t s)
{
    hipLaunchKernelGGL(k_logit, dim3((N + 3) / 4), dim3(256), 0, s, H, w, b, N, width, z);
    return hipGetLastError();
}

hipError_t launch_bce(const float* z, const float* y, int N, float pos_weight, float* dz, float* loss_sum,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_bce, dim3((N + 255) / 256), dim3(256), 0, s, z, y, N, pos_weight, dz, loss_sum);
    return hipGetLastError();
}

hipError_t launch_outer_mask(const float* dz, const float* w, const bf16* H, int N, int width, bf16* dH,
                             hipStream_t s)
{
    size_t total = (size_t)N * width;
    hipLaunchKernelGGL(k_outer_mask, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, dz, w, H, N, width, dH);
    return hipGetLastError();
}

hipError_t launch_colsum(const bf16* X, const float* sc, int N, int width, float* out, hipStream_t s)
{
    dim3 grid((width + 255) / 256, (N + kColsumRows - 1) / kColsumRows);
    hipLaunchKernelGGL(k_colsum, grid, dim3(256), 0, s, X, sc, N, width, out);
    return hipGetLastError();
}

hipError_t launch_sum(const float* x, int N, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum, dim3(64), dim3(256), 0, s, x, N, out);
    return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, bf16* pb, int n, float lr, float b1, float b2,
                       float eps, int step, hipStream_t s)
{
    float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, p, g, m, v, pb, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_to_bf16(const float* x, bf16* y, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t launch_sigmoid(float* z, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_sigmoid, dim3((n + 255) / 256), dim3(256), 0, s, z, n);
    return hipGetLastError();
}

}  // namespace vt
