// vt_build.hpp -- volume generators and the GPU octree builder.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vt_device.hpp"

namespace vt {

enum GenKind : int { GEN_BLOBS = 0, GEN_FBM = 1, GEN_SDF = 2, GEN_FOURIER = 3 };

struct GenParams {
    int kind;
    uint32_t seed;
    int64_t X, Y, Z;
    int n;                 // normalization extent (blobs: x = (i+0.5)/n; sdf: distance / n)
    int count;             // shapes / blobs / modes
    const float* shapes;   // device array of per-shape parameters
    int octaves;           // fBm
    float base_freq, gain, lacunarity, scale;
};

hipError_t launch_generate(float* out, const GenParams& p, int64_t x0, int64_t nx, hipStream_t s);
hipError_t launch_leaf_bits(const float* vol, int X, int Y, int Z, float iso, uint64_t nwords, uint64_t* bits,
                            hipStream_t s);
hipError_t launch_parent_bits(const uint64_t* child, uint64_t nwords, uint64_t* parent, hipStream_t s);
uint64_t run_blocks(uint64_t nwords);
hipError_t launch_clear_top_bit(uint64_t* word, hipStream_t s);
hipError_t launch_run_counts(const uint64_t* bits, uint64_t nwords, uint64_t nbits, uint32_t* block_sums,
                             hipStream_t s);
hipError_t launch_run_write(const uint64_t* bits, uint64_t nwords, uint64_t nbits, uint64_t code_base,
                            const uint64_t* block_offsets, uint64_t* out, hipStream_t s);

}  // namespace vt
