// vt_build.hip -- volume generators and the octree builder on the GPU.
//
// Octree builder (CreateOctree.cu:425-573 rethought for the GPU):
//   1. leaf level n: one thread per 4x4x4 Morton block of cells reads the
//      5x5x5 zero-padded samples and sets the 64 cell bits of one u64 word
//      (cell valid <=> its 8 corners are not all on one side of iso,
//      CreateOctree.cu:497-501);
//   2. inner levels: bit(node) = any of its 8 child bits (one byte of the
//      level below), CreateOctree.cu:541-571; the reference's loop skips the
//      last node of every inner level (CreateOctree.cu:550) -- reproduced
//      when `literal` is set;
//   3. run-length encoding per level (CreateOctree.cu:79-128): each word
//      finds its run starts/ends with shifts, a block-level scan plus a host
//      scan of the block totals gives output offsets, and the pairs are
//      written in code order.
// The level bitmaps double as the traversal's membership tables: the bitmap
// of level l viewed as bytes is one byte per level-(l-1) parent.
#include "vt_build.hpp"
#include "vt_internal.hpp"

#pragma clang fp contract(off)

namespace vt {

namespace {

// ---------------------------------------------------------- hashing (PRNG)
__device__ __forceinline__ uint32_t hash32(uint32_t x)
{
    // integer finalizer (xorshift-multiply), public-domain style mixing
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ float lattice_value(int ix, int iy, int iz, uint32_t seed)
{
    uint32_t h = hash32((uint32_t)ix * 0x8da6b343u ^ hash32((uint32_t)iy * 0xd8163841u ^
                 hash32((uint32_t)iz * 0xcb1ab31fu ^ seed)));
    return (float)(h >> 8) * (2.0f / 16777215.0f) - 1.0f;   // [-1, 1]
}

__device__ __forceinline__ float value_noise(float x, float y, float z, uint32_t seed)
{
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    float tx = x - fx, ty = y - fy, tz = z - fz;
    float ux = tx * tx * (3.0f - 2.0f * tx);
    float uy = ty * ty * (3.0f - 2.0f * ty);
    float uz = tz * tz * (3.0f - 2.0f * tz);
    float c000 = lattice_value(ix, iy, iz, seed), c100 = lattice_value(ix + 1, iy, iz, seed);
    float c010 = lattice_value(ix, iy + 1, iz, seed), c110 = lattice_value(ix + 1, iy + 1, iz, seed);
    float c001 = lattice_value(ix, iy, iz + 1, seed), c101 = lattice_value(ix + 1, iy, iz + 1, seed);
    float c011 = lattice_value(ix, iy + 1, iz + 1, seed), c111 = lattice_value(ix + 1, iy + 1, iz + 1, seed);
    float a = c000 + (c100 - c000) * ux, b = c010 + (c110 - c010) * ux;
    float c = c001 + (c101 - c001) * ux, d = c011 + (c111 - c011) * ux;
    float e = a + (b - a) * uy, f = c + (d - c) * uy;
    return e + (f - e) * uz;
}

__device__ __forceinline__ float fbm(float x, float y, float z, const GenParams& p)
{
    float sum = 0.0f, amp = 1.0f, norm = 0.0f, fr = p.base_freq;
    for (int o = 0; o < p.octaves; ++o) {
        sum += amp * value_noise(x * fr, y * fr, z * fr, p.seed + 7919u * (uint32_t)o);
        norm += amp;
        amp *= p.gain;
        fr *= p.lacunarity;
    }
    return p.octaves > 0 ? sum / norm * p.scale : 0.0f;   // no octave: no fBm term
}

__global__ void k_generate(float* __restrict__ out, GenParams p, int64_t x0, int64_t nx)
{
    int64_t total = nx * p.Y * p.Z;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t z = e % p.Z;
        int64_t y = (e / p.Z) % p.Y;
        int64_t x = x0 + e / (p.Z * p.Y);
        float v = 0.0f;
        if (p.kind == GEN_BLOBS) {
            float ux = ((float)x + 0.5f) / (float)p.n, uy = ((float)y + 0.5f) / (float)p.n,
                  uz = ((float)z + 0.5f) / (float)p.n;
            for (int k = 0; k < p.count; ++k) {
                const float* b = p.shapes + 5 * k;   // cx cy cz sigma amp
                float dx = ux - b[0], dy = uy - b[1], dz = uz - b[2];
                v += b[4] * expf(-(dx * dx + dy * dy + dz * dz) / (2.0f * b[3] * b[3]));
            }
        } else if (p.kind == GEN_FBM) {
            v = fbm((float)x, (float)y, (float)z, p);
        } else if (p.kind == GEN_SDF) {
            // union (min) of spheres and tori, distances in voxels, plus fBm
            float d = 3.0e38f;
            for (int k = 0; k < p.count; ++k) {
                const float* s = p.shapes + 8 * k;   // type cx cy cz r1 r2 axis pad
                float dx = (float)x - s[1], dy = (float)y - s[2], dz = (float)z - s[3];
                float dk;
                if (s[0] == 0.0f) {
                    dk = sqrtf(dx * dx + dy * dy + dz * dz) - s[4];
                } else {
                    float a0 = dx, a1 = dy, h = dz;   // torus axis selection
                    if (s[6] == 1.0f) { a0 = dx; a1 = dz; h = dy; }
                    else if (s[6] == 2.0f) { a0 = dy; a1 = dz; h = dx; }
                    float q = sqrtf(a0 * a0 + a1 * a1) - s[4];
                    dk = sqrtf(q * q + h * h) - s[5];
                }
                d = fminf(d, dk);
            }
            v = d / (float)p.n + fbm((float)x, (float)y, (float)z, p);
        } else if (p.kind == GEN_FOURIER) {
            for (int k = 0; k < p.count; ++k) {
                const float* m = p.shapes + 5 * k;   // kx ky kz phase amp (k in rad/voxel)
                v += m[4] * cosf(m[0] * (float)x + m[1] * (float)y + m[2] * (float)z + m[3]);
            }
        }
        out[e] = v;
    }
}

// -------------------------------------------------------- octree builder
__device__ __forceinline__ float vox(const float* __restrict__ v, int X, int Y, int Z, int x, int y, int z)
{
    if (x >= X || y >= Y || z >= Z) return 0.0f;   // x,y,z >= 0 always here
    return v[((size_t)x * Y + y) * Z + z];
}

// word w of the leaf bitmap = cells with Morton index [64w, 64w+64)
__global__ __launch_bounds__(256) void k_leaf_bits(const float* __restrict__ vol, int X, int Y, int Z,
                                                   float iso, uint64_t nwords, uint64_t* __restrict__ bits)
{
    uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    uint64_t m = w << 6;
    int bx = (int)compact3(m >> 2), by = (int)compact3(m >> 1), bz = (int)compact3(m);
    // 5x5x5 samples of the 4x4x4 cell block, as "below iso" flags
    uint32_t below[5][5];   // bit z of below[x][y]
#pragma unroll
    for (int x = 0; x < 5; ++x)
#pragma unroll
        for (int y = 0; y < 5; ++y) {
            uint32_t row = 0;
#pragma unroll
            for (int z = 0; z < 5; ++z) {
                float s = vox(vol, X, Y, Z, bx + x, by + y, bz + z);
                row |= ((s - iso) < 0 ? 1u : 0u) << z;
            }
            below[x][y] = row;
        }
    uint64_t out = 0;
#pragma unroll
    for (int b = 0; b < 64; ++b) {
        int cx = ((b >> 2) & 1) | (((b >> 5) & 1) << 1);
        int cy = ((b >> 1) & 1) | (((b >> 4) & 1) << 1);
        int cz = (b & 1) | (((b >> 3) & 1) << 1);
        uint32_t f = ((below[cx][cy] >> cz) & 3u) | (((below[cx][cy + 1] >> cz) & 3u) << 2) |
                     (((below[cx + 1][cy] >> cz) & 3u) << 4) | (((below[cx + 1][cy + 1] >> cz) & 3u) << 6);
        bool valid = f != 0u && f != 0xffu;   // corners on both sides of iso
        out |= (valid ? 1ull : 0ull) << b;
    }
    bits[w] = out;
}

// The same words for cubes of side >= 32, a 32^3-cell block (512 words) per
// 256-thread workgroup: its 33^3 corner samples are read once into LDS as
// below-iso flags (1.10 reads per sample, where a word alone reads 5^3
// samples for 4^3 cells), then each thread builds two words from them with
// the flags and bit layout of k_leaf_bits.
constexpr int kLeafSide = 32, kLeafS = kLeafSide + 1, kLeafZ = kLeafS + 3;   // z rows padded to 36 bytes

__global__ __launch_bounds__(256) void k_leaf_bits32(const float* __restrict__ vol, int X, int Y, int Z, float iso,
                                                     uint64_t* __restrict__ bits)
{
    __shared__ uint8_t below[kLeafS][kLeafS][kLeafZ];
    const uint64_t c0 = (uint64_t)blockIdx.x << 15;   // first cell of the block, Morton order
    const int x0 = (int)compact3(c0 >> 2), y0 = (int)compact3(c0 >> 1), z0 = (int)compact3(c0);
    // rows along z: eight 16-byte loads (z0 + 4q, 4 samples) and the 33rd
    // sample; a run that leaves the volume, or rows that are not 16-byte
    // aligned (Z not a multiple of 4, or an unaligned volume), go sample by
    // sample
    const bool vec = (Z & 3) == 0 && ((uintptr_t)vol & 15) == 0;
    for (int e = threadIdx.x; e < kLeafS * kLeafS * 9; e += 256) {
        const int row = e / 9, q = e - row * 9;
        const int x = row / kLeafS, y = row - x * kLeafS;
        const int gx = x0 + x, gy = y0 + y, gz = z0 + 4 * q;
        uint8_t* dst = &below[x][y][4 * q];
        if (q < 8 && vec && gx < X && gy < Y && gz + 3 < Z) {
            const float4 v = *reinterpret_cast<const float4*>(vol + ((size_t)gx * Y + gy) * Z + gz);
            dst[0] = (v.x - iso) < 0 ? 1 : 0;
            dst[1] = (v.y - iso) < 0 ? 1 : 0;
            dst[2] = (v.z - iso) < 0 ? 1 : 0;
            dst[3] = (v.w - iso) < 0 ? 1 : 0;
        } else {
            for (int i = 0; i < (q < 8 ? 4 : 1); ++i)
                dst[i] = (vox(vol, X, Y, Z, gx, gy, gz + i) - iso) < 0 ? 1 : 0;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 512; k += 256) {
        const uint64_t m = (uint64_t)k << 6;   // the word's first cell, relative to the block
        const int bx = (int)compact3(m >> 2), by = (int)compact3(m >> 1), bz = (int)compact3(m);
        uint32_t rows[5][5];   // bit z of rows[x][y]: sample (bx+x, by+y, bz+z) below iso
#pragma unroll
        for (int x = 0; x < 5; ++x)
#pragma unroll
            for (int y = 0; y < 5; ++y) {
                uint32_t row = 0;
#pragma unroll
                for (int z = 0; z < 5; ++z) row |= (uint32_t)below[bx + x][by + y][bz + z] << z;
                rows[x][y] = row;
            }
        uint64_t out = 0;
#pragma unroll
        for (int b = 0; b < 64; ++b) {
            int cx = ((b >> 2) & 1) | (((b >> 5) & 1) << 1);
            int cy = ((b >> 1) & 1) | (((b >> 4) & 1) << 1);
            int cz = (b & 1) | (((b >> 3) & 1) << 1);
            uint32_t f = ((rows[cx][cy] >> cz) & 3u) | (((rows[cx][cy + 1] >> cz) & 3u) << 2) |
                         (((rows[cx + 1][cy] >> cz) & 3u) << 4) | (((rows[cx + 1][cy + 1] >> cz) & 3u) << 6);
            out |= (f != 0u && f != 0xffu ? 1ull : 0ull) << b;
        }
        bits[((uint64_t)blockIdx.x << 9) + (uint64_t)k] = out;
    }
}

// level l (>= 2, 8^l >= 64) word from 8 words of level l+1
__global__ void k_parent_bits(const uint64_t* __restrict__ child, uint64_t nwords, uint64_t* __restrict__ parent)
{
    uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    uint64_t out = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        uint64_t c = child[8 * w + q];
#pragma unroll
        for (int b = 0; b < 8; ++b)
            out |= (((c >> (8 * b)) & 0xffull) != 0 ? 1ull : 0ull) << (8 * q + b);
    }
    parent[w] = out;
}

__device__ __forceinline__ void run_events(const uint64_t* __restrict__ bits, uint64_t nwords, uint64_t nbits,
                                           uint64_t w, uint64_t& starts, uint64_t& ends)
{
    uint64_t cur = bits[w];
    if ((w + 1) * 64 > nbits) cur &= (nbits - w * 64 >= 64) ? ~0ull : ((1ull << (nbits - w * 64)) - 1);
    uint64_t prev_top = (w > 0) ? (bits[w - 1] >> 63) : 0;
    uint64_t next_bot = (w + 1 < nwords && (w + 1) * 64 < nbits) ? (bits[w + 1] & 1ull) : 0;
    starts = cur & ~((cur << 1) | prev_top);
    ends = cur & ~((cur >> 1) | (next_bot << 63));
}

constexpr int kScanBlock = 256;

__global__ __launch_bounds__(kScanBlock) void k_run_counts(const uint64_t* __restrict__ bits, uint64_t nwords,
                                                           uint64_t nbits, uint32_t* __restrict__ block_sums)
{
    __shared__ uint32_t red[kScanBlock / 64];
    uint64_t w = blockIdx.x * (uint64_t)kScanBlock + threadIdx.x;
    uint32_t cnt = 0;
    if (w < nwords) {
        uint64_t s, e;
        run_events(bits, nwords, nbits, w, s, e);
        cnt = (uint32_t)(__popcll(s) + __popcll(e));
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < kScanBlock / 64; ++i) t += red[i];
        block_sums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_run_write(const uint64_t* __restrict__ bits, uint64_t nwords,
                                                          uint64_t nbits, uint64_t code_base,
                                                          const uint64_t* __restrict__ block_offsets,
                                                          uint64_t* __restrict__ out)
{
    __shared__ uint32_t scan[kScanBlock];
    uint64_t w = blockIdx.x * (uint64_t)kScanBlock + threadIdx.x;
    uint64_t s = 0, e = 0;
    if (w < nwords) run_events(bits, nwords, nbits, w, s, e);
    uint32_t cnt = (uint32_t)(__popcll(s) + __popcll(e));
    scan[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {   // inclusive Hillis-Steele
        uint32_t v = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 0;
        __syncthreads();
        scan[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t pos = block_offsets[blockIdx.x] + scan[threadIdx.x] - cnt;
    uint64_t ev = s | e;
    while (ev) {
        int b = __ffsll((long long)ev) - 1;
        uint64_t code = code_base + w * 64 + (uint64_t)b;
        if ((s >> b) & 1) out[pos++] = code;
        if ((e >> b) & 1) out[pos++] = code;
        ev &= ev - 1;
    }
}

__global__ void k_clear_top_bit(uint64_t* w) { *w &= ~(1ull << 63); }

__global__ __launch_bounds__(256) void k_node_minmax(const float* __restrict__ vol, int X, int Y, int Z, int n,
                                                     int level, float* __restrict__ mn, float* __restrict__ mx)
{
    __shared__ float smin[4], smax[4];
    uint64_t idx = blockIdx.x;   // node index within the level
    uint64_t code = (1ull << (3 * level)) | idx;
    int x0, y0, z0;
    node_min(code, level, n, x0, y0, z0);
    const int s1 = (1 << (n - level)) + 1;
    const int total = s1 * s1 * s1;
    float lo = 3.0e38f, hi = -3.0e38f;
    for (int e = threadIdx.x; e < total; e += 256) {
        int z = e % s1, y = (e / s1) % s1, x = e / (s1 * s1);
        float v = vox(vol, X, Y, Z, x0 + x, y0 + y, z0 + z);
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, off));
        hi = fmaxf(hi, __shfl_xor(hi, off));
    }
    if ((threadIdx.x & 63) == 0) { smin[threadIdx.x >> 6] = lo; smax[threadIdx.x >> 6] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        mn[idx] = fminf(fminf(smin[0], smin[1]), fminf(smin[2], smin[3]));
        mx[idx] = fmaxf(fmaxf(smax[0], smax[1]), fmaxf(smax[2], smax[3]));
    }
}

}  // namespace

hipError_t launch_node_minmax(const float* vol, int X, int Y, int Z, int n, int level, float* mn, float* mx,
                              hipStream_t s)
{
    hipLaunchKernelGGL(k_node_minmax, dim3((unsigned)(1ull << (3 * level))), dim3(256), 0, s, vol, X, Y, Z, n, level,
                       mn, mx);
    return hipGetLastError();
}

hipError_t launch_clear_top_bit(uint64_t* word, hipStream_t s)
{
    hipLaunchKernelGGL(k_clear_top_bit, dim3(1), dim3(1), 0, s, word);
    return hipGetLastError();
}

hipError_t launch_generate(float* out, const GenParams& p, int64_t x0, int64_t nx, hipStream_t s)
{
    int64_t total = nx * p.Y * p.Z;
    int64_t blocks = std::min<int64_t>((total + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_generate, dim3((unsigned)blocks), dim3(256), 0, s, out, p, x0, nx);
    return hipGetLastError();
}

hipError_t launch_leaf_bits(const float* vol, int X, int Y, int Z, float iso, uint64_t nwords, uint64_t* bits,
                            hipStream_t s)
{
    if (nwords >= 512) {   // a cube of side >= 32: whole 32^3 blocks
        hipLaunchKernelGGL(k_leaf_bits32, dim3((unsigned)(nwords / 512)), dim3(256), 0, s, vol, X, Y, Z, iso, bits);
    } else {
        hipLaunchKernelGGL(k_leaf_bits, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s, vol, X, Y, Z, iso,
                           nwords, bits);
    }
    return hipGetLastError();
}

hipError_t launch_parent_bits(const uint64_t* child, uint64_t nwords, uint64_t* parent, hipStream_t s)
{
    hipLaunchKernelGGL(k_parent_bits, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s, child, nwords,
                       parent);
    return hipGetLastError();
}

uint64_t run_blocks(uint64_t nwords) { return (nwords + kScanBlock - 1) / kScanBlock; }

hipError_t launch_run_counts(const uint64_t* bits, uint64_t nwords, uint64_t nbits, uint32_t* block_sums,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_run_counts, dim3((unsigned)run_blocks(nwords)), dim3(kScanBlock), 0, s, bits, nwords,
                       nbits, block_sums);
    return hipGetLastError();
}

hipError_t launch_run_write(const uint64_t* bits, uint64_t nwords, uint64_t nbits, uint64_t code_base,
                            const uint64_t* block_offsets, uint64_t* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_run_write, dim3((unsigned)run_blocks(nwords)), dim3(kScanBlock), 0, s, bits, nwords,
                       nbits, code_base, block_offsets, out);
    return hipGetLastError();
}

}  // namespace vt
