// vt_internal.hpp -- shared host/device declarations of the HIP library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "vt_device.hpp"

namespace vt {

struct HitRecord {        // 32 bytes, one per ray
    uint64_t node;        // levelO node code of the hit, 0 on miss
    uint64_t brick;       // levelC brick code, 0 on miss
    int32_t voxel[3];     // floor of the secant point
    float depth;          // |Xnew - origin|
};
static_assert(sizeof(HitRecord) == 32, "hit record layout");

// the brick pool as seen by the march
struct BrickView {
    const int32_t* brick_slot;   // brick index -> slot, -1 when absent
    const float* pool;
    int S, inc;
    size_t slot_elems;
    int level_c, shift;          // shift = 3*(levelO - levelC)
    uint64_t brick_base;         // 8^levelC
    const float2* submm;         // per slot G^3 sub-block (min, max), or null
    int G;                       // sub-blocks per axis (cubeDim / 4)
    // per slot, the side of iso each sub-block's (min, max) proves, 2 bits a
    // sub-block (16 to a word, sub-block sb in word sb >> 4), computed for
    // this pass's iso; null: decide from submm
    const uint32_t* sides = nullptr;
    int side_words = 0;          // words per slot
};

struct Shading {
    float iso, lx, ly, lz, color_scale;
};

enum WalkResult : int { WALK_MISS = 0, WALK_HIT = 1, WALK_ABSENT = 2 };

// Parked-ray phase byte: bits 0-4 = entries in the ray's deferral list,
// bit 7 = the output already holds the ray's tentative result (the walk
// ended past its deferred nodes; a deferred node that crosses overrides it).
enum RayPhase : uint8_t { PHASE_NMASK = 0x1f, PHASE_TENT = 0x80 };
constexpr int kMaxDefer = 16;

// per-frame device counters (one 64-byte block)
struct FrameCounters {
    uint32_t req_count;       // distinct bricks requested in this pass
    uint32_t parked_count;    // rays parked waiting for a brick
    uint32_t err;             // bit 0: a sample fell outside its brick slot
    uint32_t pad0;
    unsigned long long samples;   // trilinear samples taken (march + normal)
    unsigned long long nodes;     // nodes marched
    uint32_t pad[8];
};

// One frame of a multi-frame pass (vt_frames): its camera, shading and
// outputs.  Rays of several frames share the launches, the per-ray state
// arrays (ray id = frame slot << rshift | pixel) and the brick cache.
struct RenderJob {
    Camera cam;
    Shading sh;
    float4* rgba;
    HitRecord* hits;
};
constexpr int kMaxFrames = 32;

struct RenderArgs {
    Camera cam;
    OctreeView oct;
    int level_c;
    int shift;                   // 3*(levelO - levelC)
    uint64_t brick_base;         // 8^levelC
    const int32_t* brick_slot;   // brick index -> slot, -1 when absent
    const float* pool;           // slot-major brick pool
    int S;                       // slot side = cubeDim + 2*inc
    int inc;
    size_t slot_elems;           // S^3
    float iso, lx, ly, lz, color_scale;
    // ray state
    uint64_t* ray_node;          // resume node of a parked ray
    uint8_t* ray_phase;
    uint64_t* ray_defer;         // [max_defer][R] deferred nodes, or null
    int max_defer;               // 0: park at the first absent brick
    int node_budget;             // walk nodes a ray may take in this pass before it parks (0: no limit)
    int supertile;               // first pass: S x S blocks of 8x8 tiles dealt to the XCDs (<= 1: single tiles)
    uint32_t* next;              // persistent form: 8 work counters (one per XCD part), zero at launch
    const uint32_t* in_list;     // null: first pass over the whole frame
    uint32_t in_count;
    uint32_t* out_list;
    uint32_t* req_bits;
    uint32_t* req_list;          // host-visible; entry = brick index + 1
    uint8_t* slot_ref;
    uint8_t* brick_trace;        // optional, 1 byte per brick
    const float2* submm;         // sub-block min/max per slot, or null (no skipping)
    int G;
    const uint32_t* sides;       // per slot proven sides for iso side_iso (see BrickView), or null
    int side_words;
    float side_iso;              // rays whose iso differs use submm
    FrameCounters* ctr;
    float4* rgba;
    HitRecord* hits;
    int tiles_x;                 // 16x16 pixel tiles per row
    int tiles_total;
    int collect_stats;
    // multi-frame passes (jobs != null): in_list holds the parked rays of
    // frames already in flight (their later passes); the first pass of the
    // nfresh frames fresh_slot[] follows in the same launch, tile by tile
    const RenderJob* jobs;
    int rshift;                  // ray id = frame slot << rshift | pixel
    uint32_t defer_stride;       // rays per deferral list (frame slots << rshift)
    int max_defer_late;          // deferral of the resumed rays (max_defer: of the first passes)
    uint32_t* job_parked;        // per frame slot: rays parked in this launch (their count in its list)
    // in_list/out_list hold one list per frame slot, at slot << rshift; the
    // in_list groups of 64 entries of frame s are groups
    // [list_block0[s], list_block0[s + 1]) of the launch
    uint32_t list_count[kMaxFrames];
    uint32_t list_block0[kMaxFrames + 1];
    int nfresh;
    uint32_t fresh_slot[kMaxFrames];
    uint32_t fresh_tile0[kMaxFrames + 1];   // first 8x8 tile of each fresh frame in the launch (prefix sums)
};

struct TraceArgs {
    Camera cam;
    OctreeView oct;
    int max_nodes;
    int cap_total;
    uint64_t* nodes;             // [R * max_nodes]
    int32_t* counts;             // [R]
};

// visibleCube_t of the reference (config.hpp:13-19), 32 bytes on x86-64
struct VisibleCube {
    uint64_t id;
    float* data;
    uint8_t state;
    uint8_t pad[7];
    uint64_t cubeID;
};
static_assert(sizeof(VisibleCube) == 32, "visibleCube_t layout");

enum CubeState : uint8_t { ST_NOCUBE = 0, ST_NOCACHED = 1, ST_CACHED = 2, ST_PAINTED = 4, ST_CUBE = 8 };
enum TravState : uint8_t { TRAV_FRESH = 0, TRAV_ACTIVE = 1, TRAV_EXHAUSTED = 2 };

struct StageTraverseArgs {
    Camera cam;
    OctreeView oct;
    VisibleCube* cubes;
    uint8_t* trav;            // per-ray traversal state (the reference's stackActual)
};

struct StageRaycastArgs {
    Camera cam;
    int n, level_o, level_c, inc, S;
    float iso, lx, ly, lz, color_scale;
    VisibleCube* cubes;
    float4* rgba;
    uint32_t* err;
};

struct CubeArgs {
    Camera cam;
    float iso, lx, ly, lz;
    const float* cube;        // (dim+2*inc)^3 floats, z fastest
    int mnx, mny, mnz;        // interior min corner (voxels)
    int dim, inc;
    float4* rgba;
    uint32_t* err;
};

// one admitted render request of the serving loop
struct ServeJob {
    Camera cam;
    OctreeView oct;      // membership tables for this request's isovalue
    BrickView bv;
    Shading sh;
    float4* rgba;
    HitRecord* hits;
    int tiles_x;
};

struct ServeTile {
    uint32_t job;
    uint32_t tile;
};

struct ServeCounters {
    uint32_t nwork;      // tiles in this launch
    uint32_t err;
};

// launchers (vt_kernels.hip)
// the serving pass over ctr->nwork tiles (of at most max_tiles), 2^lg lanes per ray
hipError_t launch_serve_pass(const ServeJob* jobs, const ServeTile* work, ServeCounters* ctr, int max_tiles, int lg,
                             hipStream_t s);
// vt_build.hip: min and max over the (s+1)^3 zero-padded samples of every
// node of one level (s = 2^(n-level)); one workgroup per node
hipError_t launch_node_minmax(const float* vol, int X, int Y, int Z, int n, int level, float* mn, float* mx,
                              hipStream_t s);
hipError_t launch_render_cube(const CubeArgs& a, hipStream_t s);
hipError_t launch_stage_traverse(const StageTraverseArgs& a, hipStream_t s);
hipError_t launch_stage_raycast(const StageRaycastArgs& a, hipStream_t s);
hipError_t launch_render_persistent(const RenderArgs& a, int max_workgroups, hipStream_t s);
int render_persistent_workgroups(int device);   // resident workgroups of the persistent pass on `device`
hipError_t launch_render_pass(const RenderArgs& a, hipStream_t s);
// a pass with 2^lg lanes per ray, lg in 0..6 (first pass: in_list null)
hipError_t launch_render_wide(const RenderArgs& a, int lg, hipStream_t s);
// multi-frame pass: the in_list rays (one lane each), then the tiles of the
// fresh frames; with no fresh frame and lg > 0, the in_list rays with 2^lg lanes each
hipError_t launch_render_multi(const RenderArgs& a, int lg, hipStream_t s);
hipError_t launch_trace_nodes(const TraceArgs& a, hipStream_t s);
// Assemble bricks with halo from a device volume [X][Y][Z] (zero outside).
hipError_t launch_bricks_from_volume(const float* vol, int X, int Y, int Z,
                                     const int64_t* brick_codes, const int32_t* slots, int count,
                                     int level_c, int n, int cube_dim, int inc, size_t stride,
                                     float* pool, hipStream_t s);   // slot b at pool + b * stride
// (min, max) of every 4x4x4-cell sub-block (5^3 samples) of the given slots
hipError_t launch_submm(const float* pool, const int32_t* slots, int count, int S, int inc, int G, size_t stride,
                        float2* submm, hipStream_t s);
// proven-side words of `count` slots (slots[i], or slot i when slots is null)
// from their sub-block (min, max) for `iso`
hipError_t launch_sides(const float2* submm, const int32_t* slots, int count, int G, float iso, uint32_t* sides,
                        int side_words, hipStream_t s);
// Copy bricks of slot_elems floats from (device-visible) host memory into
// the pool, slot b at pool + b * stride, with `split` workgroups per brick
// (0: the default of 8).
hipError_t launch_copy_bricks(const float* const* src, const int32_t* dst_slots, int count,
                              size_t slot_elems, size_t stride, float* pool, hipStream_t s, int split = 0);
// table[idx[i]] = val[i]
hipError_t launch_scatter_i32(int32_t* table, const int64_t* idx, const int32_t* val, int count,
                              hipStream_t s);
// Per-pass state that the service between two render passes resets: the
// request bitmap, the first `nctr` counter words (request and parked counts)
// and the per-slot reference bytes (already read back by the host).
struct PassReset {
    uint32_t* req_bits;
    int64_t nwords;
    uint32_t* ctr;
    int nctr;
    uint8_t* ref;
    int64_t nref;
};
// table[idx[i]] = val[i] for i < count, then the resets of `r`: one launch
// ends the service between two passes
hipError_t launch_finish_service(int32_t* table, const int64_t* idx, const int32_t* val, int count,
                                 const PassReset& r, hipStream_t s);
// streaming copy of `bytes` (multiple of 16) used to calibrate HBM counters
hipError_t launch_copy_probe(const float* src, float* dst, size_t bytes, hipStream_t s);
// camera rays (Camera.cu:39-74)
hipError_t launch_camera_rays(const Camera& cam, float* dirs, hipStream_t s);

}  // namespace vt
