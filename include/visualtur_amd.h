/*
 * visualtur_amd.h -- C ABI of the MI355X isosurface ray-casting library.
 *
 * The reference (carlosduelo/visualTur) exposes this path as C++ classes
 * whose methods launch CUDA kernels directly.  Here the same path sits
 * behind plain C entry points (opaque handles, plain pointers and sizes,
 * int status codes, no exceptions); the C++ classes in visualTur.hpp,
 * Camera.hpp, Octree.hpp, lruCache.hpp, rayCaster.hpp and FileManager.hpp
 * are written on top of these.  Each declaration names the reference
 * interface it replaces.
 *
 * Conventions
 *   - every function returns VT_OK (0) or a negative VT_E_* status;
 *     vt_last_error() returns the message of the last failure on the
 *     calling thread;
 *   - pointers named d_* are device (HBM) pointers, e.g. from vt_malloc;
 *   - one vt_context per host thread and device; calls on one context are
 *     not re-entrant.  Its kernels run on the context's own HIP stream.
 */
#ifndef VISUALTUR_AMD_H
#define VISUALTUR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VT_OK = 0,
    VT_E_ARG = -1,      /* invalid argument or size */
    VT_E_HIP = -2,      /* HIP runtime error */
    VT_E_IO = -3,       /* file could not be opened / read / written */
    VT_E_FORMAT = -4,   /* bad octree magic or malformed run list */
    VT_E_STATE = -5,    /* call out of order (e.g. no octree uploaded) */
    VT_E_NOMEM = -6,    /* allocation failed */
    VT_E_RANGE = -7     /* a sample fell outside its brick slot */
};

typedef struct vt_context vt_context;
typedef struct vt_file vt_file;

/* Camera pose and screen, world units = voxels.  Same fields as the
 * reference Camera+Screen state (Camera.hpp:31-41, Screen.hpp:11-18);
 * h and w are the screen plane size 2*d*tan(fov) (Screen.cpp:18-19). */
typedef struct {
    float pos[3];
    float look[3];
    float up[3];
    float right[3];
    float h, w, distance;
    int32_t W, H;
} vt_camera;

/* Per-frame constants the reference spreads over visualTur/rayCaster
 * (visualTur.cu:30, rayCaster.cu:245, visualTur.hpp:24-27). */
typedef struct {
    int32_t levelO;        /* octreeLevel: level the traversal emits */
    int32_t levelC;        /* levelCubes: brick level of the cache */
    float iso;             /* isovalue (the reference takes it from the octree file) */
    float light[3];        /* light position, (0,512,0) in the reference */
    float color_scale;     /* 256 in the reference */
} vt_render_params;

/* Hit record, 32 bytes per ray (the reference keeps only RGBA; these are
 * the node/brick/voxel of visibleCube_t at the moment of the hit,
 * config.hpp:13-19). */
typedef struct {
    uint64_t node;         /* levelO node Morton code, 0 on miss */
    uint64_t brick;        /* levelC brick code, 0 on miss */
    int32_t voxel[3];      /* floor of the secant point */
    float depth;           /* |Xnew - origin| */
} vt_hit;

typedef struct {
    int32_t passes;             /* kernel passes (the reference's while(notEnd) iterations) */
    int32_t err;                /* nonzero: a sample fell outside its slot */
    int64_t bricks_requested;   /* distinct bricks missing per pass, summed */
    int64_t bricks_uploaded;    /* bricks copied into the device pool */
    int64_t host_tier_misses;   /* bricks read from the file during the frame */
    int64_t rays_parked;        /* ray-waits summed over passes */
    uint64_t samples;           /* trilinear samples (when stats enabled) */
    uint64_t nodes;             /* nodes marched (when stats enabled) */
    double ms_total;            /* frame wall time */
    double ms_kernel;           /* device time of the render passes */
    double ms_service;          /* host time spent servicing misses */
    int64_t bricks_early;       /* of bricks_uploaded: copied while the first pass ran */
} vt_frame_stats;

typedef struct {
    int64_t gpu_slots, gpu_resident;
    int64_t cpu_slots, cpu_resident;
    int64_t lookups_missed;     /* device misses (bricks requested) */
    int64_t uploads;            /* device pool fills */
    int64_t file_reads;         /* host tier fills from the file */
    int64_t evictions;
} vt_cache_stats;

/* ------------------------------------------------------------ context */

/* Replaces the implicit CUDA context + cudaSetDevice (testVisualTur.cu:18-23);
 * device -1 = the calling thread's current device. */
int vt_context_create(int device, vt_context **out);
int vt_context_destroy(vt_context *ctx);
/* Message of the last failure on this thread (ctx may be NULL). */
int vt_last_error(vt_context *ctx, char *buf, size_t len);
/* Replaces cudaDeviceSynchronize after each launch (rayCaster.cu:382). */
int vt_synchronize(vt_context *ctx);

/* Device memory helpers (the reference calls cudaMalloc/cudaMemcpy inline,
 * e.g. visualTur.cu:17, testVisualTur.cu:70,90). */
int vt_malloc(vt_context *ctx, size_t bytes, void **out);
int vt_free(vt_context *ctx, void *p);
int vt_memcpy_h2d(vt_context *ctx, void *d_dst, const void *src, size_t bytes);
int vt_memcpy_d2h(vt_context *ctx, void *dst, const void *d_src, size_t bytes);
int vt_memset(vt_context *ctx, void *d_dst, int value, size_t bytes);
/* Free and total device memory of `device` (the cudaMemGetInfo that sizes
 * the viewer's brick cache, drawPixel.cu:214-217).  Like the reference's
 * cudaSetDevice before it (drawPixel.cu:205), it makes `device` the calling
 * thread's current device, which vt_context_create(-1, ...) then uses. */
int vt_device_memory(int32_t device, size_t *free_bytes, size_t *total_bytes);

/* -------------------------------------------------------------- octree */

/* Replaces the upload half of Octree::Octree (Octree.cu:820-848): run lists
 * of levels 0..maxLevel, concatenated, sizes[l] u64 entries each (sorted
 * inclusive [start,end] pairs).  Malformed lists are rejected. */
int vt_octree_upload(vt_context *ctx, const uint64_t *runs, const int32_t *sizes,
                     int32_t nLevels, int32_t maxLevel);
/* Replaces Octree::Octree(file, camera, maxLevel) (Octree.cu:769-863):
 * reads the reference binary file, uploads levels 0..maxLevel, returns the
 * file's isovalue (Octree::getIsosurface, Octree.cu:872) and nLevels
 * (Octree::getnLevels, Octree.cu:877). */
int vt_octree_load_file(vt_context *ctx, const char *path, int32_t maxLevel,
                        float *iso, int32_t *nLevels, int32_t realDim[3]);

/* -------------------------------------------------------------- camera */

/* Replaces Camera::UpdateRays / cuda_updateRays (Camera.cu:39-83): W*H
 * normalized float3 directions into d_dirs (pixel id = i*W + j). */
int vt_camera_rays(vt_context *ctx, const vt_camera *cam, float *d_dirs);
/* Screen::Screen plane size (Screen.cpp:18-19): h = 2 d tan(fovH deg),
 * w = 2 d tan(fovW deg) -- the fov arguments are half-angles. */
int vt_screen_size(float distance, float fovH, float fovW, float *h, float *w);

/* --------------------------------------------------------- brick cache */

/* Replaces lruCache::lruCache (lruCache.cu:101-129): a device pool of
 * gpuSlots bricks of (cubeDim+2*cubeInc)^3 floats at level levelC, backed
 * by a pinned host tier of cpuSlots bricks (0 = no host tier). */
int vt_cache_create(vt_context *ctx, int32_t levelC, int32_t cubeDim, int32_t cubeInc,
                    int64_t gpuSlots, int64_t cpuSlots);
/* Brick source = HDF5 dataset (lruCache.cu:118 opens a FileManager). */
int vt_cache_attach_hdf5(vt_context *ctx, const char *file, const char *dataset);
/* Brick source = a brick-major file written by vt_volume_write_bricks_hdf5:
 * each brick one record (no reference counterpart: the reference reads
 * every brick as a box of the volume dataset, lruCache.cu:226-266 through
 * FileManager.cu:86-167; the records hold exactly those boxes).  Its
 * levelC/cubeDim/cubeInc must be the cache's; a brick the octree needs that
 * the file lacks is VT_E_ARG. */
int vt_cache_attach_hdf5_bricks(vt_context *ctx, const char *file, const char *dataset);
/* Brick source = a volume already in HBM, C order [X][Y][Z] (misses are
 * assembled on the device; no host tier involved). */
int vt_cache_attach_device_volume(vt_context *ctx, const float *d_vol, int32_t X, int32_t Y,
                                  int32_t Z);
/* Brick source = a pinned host tier filled once from a volume in HBM (C
 * order [X][Y][Z]): every brick holding octree nodes (the octree must be
 * uploaded down to levelC; at most cpuSlots of them, VT_E_ARG otherwise) is
 * assembled on the device with its zero-padded halo and copied to pinned
 * host memory; device misses then cross the host link as from a tier filled
 * from HDF5 (lruCache.cu:226-266 without the file).  The tier holds exactly
 * the bricks with nodes, the only ones a ray can request. */
int vt_cache_attach_host_from_device(vt_context *ctx, const float *d_vol, int32_t X, int32_t Y, int32_t Z);
/* Start filling the host tier from the file on the reader thread
 * (wait=1 blocks until done). */
int vt_cache_prefetch_host(vt_context *ctx, int wait);
/* Load the given bricks (level-levelC codes) into the device pool ahead of
 * use, as far as capacity allows (lruCache::updateCube, lruCache.cu:168-296). */
int vt_cache_load_bricks(vt_context *ctx, const uint64_t *codes, int64_t count);
/* Copy the device slot of a resident brick (level-levelC code) to out[S^3],
 * S = cubeDim + 2*cubeInc, x slowest: the box [min - cubeInc, min + cubeDim +
 * cubeInc) of the volume, zero outside it, as lruCache's GPU slot holds it
 * (the box read into a CPU slot, lruCache.cu:226 with FileManager.cu:86-167,
 * then copied to the GPU slot, lruCache.cu:266).  VT_E_STATE when the brick
 * is not resident. */
int vt_cache_read_brick(vt_context *ctx, uint64_t code, float *out);
/* Drop every device-resident brick. */
int vt_cache_clear(vt_context *ctx);
/* lruCache::printStatistics (lruCache.cu:318-322) as numbers. */
int vt_cache_get_stats(vt_context *ctx, vt_cache_stats *out);
/* Replaces lruCache::changeDimensionCube (lruCache.cu:141-166). */
int vt_cache_destroy(vt_context *ctx);

/* -------------------------------------------------------------- render */

/* Replaces visualTur::updateVisibleCubes (visualTur.cu:167-228): renders
 * one full frame into d_rgba (W*H float4, pixel id = i*W + j, caller-owned
 * device memory as in testVisualTur.cu:70) and, if non-NULL, d_hits (W*H
 * vt_hit).  Runs traversal+march+shade passes until every ray is painted,
 * loading missing bricks between passes.  stats may be NULL. */
int vt_render_frame(vt_context *ctx, const vt_camera *cam, const vt_render_params *p,
                    float *d_rgba, vt_hit *d_hits, vt_frame_stats *stats);
/* vt_render_frame for callers that cannot allocate device memory (code
 * that must not call hipMalloc: tight serving loops, harnesses that forbid
 * it).  No counterpart in the reference.  The per-ray state of the render
 * passes lives in a file-scope device array of the library (66 MiB on every
 * GPU it runs on, held until the process ends) instead of buffers the
 * context grows; the brick cache and its staging are allocated beforehand
 * by vt_cache_create.  Limits: at most 2097152 rays (1920x1080 fits),
 * max_defer and max_defer_late <= 2, d_hits required, trace recording off
 * (VT_E_ARG / VT_E_STATE otherwise, before any device work).  Calls on one
 * GPU share the array; within one process they are serialized. */
int vt_render_frame_noalloc(vt_context *ctx, const vt_camera *cam, const vt_render_params *p,
                            float *d_rgba, vt_hit *d_hits, vt_frame_stats *stats);
/* Enable per-frame sample/node counters and the visited-brick trace. */
int vt_set_frame_options(vt_context *ctx, int collect_stats, int record_trace);
/* Frame options with no reference counterpart; results are identical for
 * every setting.  skip_empty=1 (default) skips the evaluation of samples
 * whose 4x4x4-cell sub-block lies strictly on one side of iso.  max_defer
 * (0..16, default 1) is how many nodes with an absent brick a ray may walk
 * past in the first pass (their bricks are requested, and they are marched
 * in a later pass before the ray's tentative result is accepted);
 * max_defer_late (0..16, default 2) is the same for the later passes, which carry only
 * the parked rays; 0 parks a ray at its first absent brick, as the
 * reference's staged loop does (visualTur.cu:175-220).  persistent=1 runs each pass as a fixed grid whose
 * lanes take a new ray when theirs finishes; 0 (default) runs one lane per
 * ray in one-wave workgroups of 8x8 pixels. */
int vt_set_march_options(vt_context *ctx, int skip_empty, int max_defer, int max_defer_late, int persistent);
/* Wide march in the passes after the first, which carry only the parked
 * rays (no reference counterpart; results are identical for every
 * setting): each ray gets 2^k lanes (k <= 6) that march its nodes together,
 * the largest 2^k with rays x 2^k <= lane_budget (default 262144, the lanes
 * of about one wave per SIMD); 0 keeps one lane per ray in
 * every pass; -k (k in 1..6) gives every later pass 2^k lanes per ray.
 * first_pass_budget does the same for the first pass, which carries every
 * ray of the frame (default 2097152: 8 lanes per ray at 512x512, one lane
 * per ray from 1920x1080 on).  Not used with persistent=1. */
int vt_set_wide_march(vt_context *ctx, int64_t lane_budget, int64_t first_pass_budget);
/* Loads beside the first pass (no reference counterpart; results are
 * identical for every setting): before the first pass of a frame, clock
 * victims are evicted until min(reserve, bricks the previous frame's first
 * pass requested) slots are free; while the pass runs, the bricks it
 * requests are copied into free slots, and they become visible from the
 * second pass on (page-table entries are written between passes only).
 * Default 256 (about what the host link moves during a 1080p first pass);
 * 0 loads only between passes. */
int vt_set_early_loads(vt_context *ctx, int64_t reserve);
/* Walk budget of the first pass (no reference counterpart; results are
 * identical for every setting): a ray that has taken `nodes` walk nodes in
 * the first pass parks on its next node without requesting anything, and
 * finishes in the later passes (with the wide march).  Default 16; 0: no
 * limit. */
int vt_set_first_pass_budget(vt_context *ctx, int32_t nodes);
/* Tile order of the one-lane first pass, of single frames and of frames in
 * flight (no reference counterpart; results are identical for every
 * setting): supertile S > 1 deals S x S blocks of 8x8-pixel tiles
 * round-robin to the XCDs, so tiles that share bricks share one L2
 * (default 4); 1 deals single tiles.  S in [1, 16]. */
int vt_set_first_pass_tiles(vt_context *ctx, int32_t supertile);
/* Visited-brick trace of the last frame: out[i] = 1 if brick
 * (i + 8^levelC) was marched by at least one ray.  n = 8^levelC. */
int vt_get_brick_trace(vt_context *ctx, uint8_t *out, int64_t n);

/* Octree::getBoxIntersected repeated to exhaustion (Octree.cu:883-901):
 * for every pixel, the first maxNodes level-finalLevel nodes in emission
 * order into d_nodes[id*maxNodes + k], and the emission count (capped at
 * capTotal) into d_counts[id]. */
int vt_trace_nodes(vt_context *ctx, const vt_camera *cam, int32_t finalLevel,
                   int32_t maxNodes, int32_t capTotal, uint64_t *d_nodes, int32_t *d_counts);

/* ------------------------------------------------------ staged protocol */

/* visibleCube_t of the reference (config.hpp:13-19), 32 bytes. */
typedef struct {
    uint64_t id;           /* current node (levelO Morton code) */
    float *data;           /* device pointer of the brick slot (set by vt_stage_cache) */
    uint8_t state;         /* VT_NOCUBE / VT_NOCACHED / VT_CACHED / VT_PAINTED / VT_CUBE */
    uint8_t pad_[7];
    uint64_t cubeID;       /* brick code at levelC */
} vt_visible_cube;
enum { VT_NOCUBE = 0, VT_NOCACHED = 1, VT_CACHED = 2, VT_PAINTED = 4, VT_CUBE = 8 };

/* Octree::resetState (Octree.cu:903-911): every ray restarts at the root. */
int vt_stage_reset(vt_context *ctx, int64_t numRays);
/* Octree::getBoxIntersected (Octree.cu:883-901) on device cubes: each
 * NOCUBE ray advances to its next front-to-back node of level finalLevel
 * (state CUBE) or is exhausted (id 0). */
int vt_stage_traverse(vt_context *ctx, const vt_camera *cam, int32_t finalLevel, vt_visible_cube *d_cubes);
/* lruCache::updateCache (lruCache.cu:298-313) on host cubes: rays in CUBE or
 * NOCACHED get their brick (CACHED, data = slot) or stay NOCACHED when the
 * pool is full of bricks this same pass uses.  levelO = octreeLevel,
 * nLevels = the octree depth. */
int vt_stage_cache(vt_context *ctx, vt_visible_cube *h_cubes, int64_t num, int32_t levelO, int32_t nLevels);
/* rayCaster::renderCube / cuda_rayCasterCube (rayCaster.cu:267-357, 385-394):
 * every pixel marched through one brick d_cube of (cubeDim+2*cubeInc)^3
 * floats whose interior starts at voxel minBox (used by testRayCaster). */
int vt_render_cube(vt_context *ctx, const vt_camera *cam, float iso, const float light[3],
                   const float *d_cube, const int32_t minBox[3], const int32_t cubeDim[3], int32_t cubeInc,
                   float *d_rgba);
/* rayCaster::render (rayCaster.cu:375-384): NOCUBE rays are painted white;
 * CACHED rays march their node in cube->data and become PAINTED (hit) or
 * NOCUBE (no crossing). */
int vt_stage_raycast(vt_context *ctx, const vt_camera *cam, const vt_render_params *p, int32_t nLevels,
                     int32_t cubeDim, int32_t cubeInc, vt_visible_cube *d_cubes, float *d_rgba);

/* ------------------------------------------------------- octree builder */

/* Replaces OctreeM::OctreeM + createTree (CreateOctree.cu:300-334,
 * 425-573) for a volume already in HBM (C order [X][Y][Z], zero outside):
 * builds every level 0..nLevels on the device, keeps their run lists in the
 * context, and installs levels 0..maxLevel as the traversal's octree.
 * literal=1 reproduces the reference loop that never visits the last node
 * of an inner level (CreateOctree.cu:550).  sizes_out (may be NULL) gets
 * nLevels+1 run-list sizes (u64 entries per level); needs 2 <= nLevels <= 12. */
int vt_octree_build(vt_context *ctx, const float *d_vol, int32_t X, int32_t Y, int32_t Z, float iso,
                    int32_t literal, int32_t maxLevel, int32_t *nLevels_out, int64_t *sizes_out);
/* Copy the run list of one level of the last build to host memory. */
int vt_octree_get_runs(vt_context *ctx, int32_t level, uint64_t *out, int64_t count);
/* OctreeM::writeToFile (CreateOctree.cu:342-371) for the last build. */
int vt_octree_write_file(vt_context *ctx, const char *path);

/* ---------------------------------------------------- volume generators */

/* Seeded synthetic volumes of the benchmark configurations (no reference
 * counterpart: the reference reads user-supplied HDF5 files). */
enum { VT_GEN_BLOBS = 0, VT_GEN_FBM = 1, VT_GEN_SDF = 2, VT_GEN_FOURIER = 3 };
typedef struct {
    int32_t kind;
    uint32_t seed;
    int32_t n;              /* normalization extent */
    int32_t count;          /* number of blobs / shapes / modes */
    const float *shapes;    /* host array: blobs 5 floats (cx cy cz sigma amp, unit cube);
                               sdf 8 floats (type cx cy cz r1 r2 axis 0, voxels);
                               fourier 5 floats (kx ky kz phase amp, rad/voxel) */
    int32_t octaves;        /* fBm part (fbm and sdf kinds) */
    float base_freq, gain, lacunarity, scale;
} vt_gen_params;
/* Generate x-slab [x0, x0+nx) of an X*Y*Z volume into d_out (nx*Y*Z floats). */
int vt_volume_generate(vt_context *ctx, const vt_gen_params *p, int64_t X, int64_t Y, int64_t Z,
                       int64_t x0, int64_t nx, float *d_out);
/* Write a device volume to an HDF5 dataset, x-slab by x-slab
 * (FileManager::CreateFile semantics, FileManager.cu:169-243); chunk 0 =
 * contiguous. */
/* Write the bricks `codes` (ascending Morton codes at levelC) of a device
 * volume as a brick-major HDF5 file: dataset `dataset` of shape (K, S, S, S),
 * S = cubeDim + 2 cubeInc, record k = the zero-padded box [min - cubeInc,
 * min - cubeInc + S) of brick codes[k] (what FileManager's box read returns
 * for it, FileManager.cu:86-167), one record per chunk; `dataset`_codes
 * holds the codes and attributes hold levelC, cubeDim, cubeInc and the
 * volume dims.  (The configuration's "32^3 chunks with a halo": a brick is
 * then one contiguous read instead of a box over up to eight chunks.) */
int vt_volume_write_bricks_hdf5(vt_context *ctx, const float *d_vol, int64_t X, int64_t Y, int64_t Z, int32_t levelC,
                                int32_t cubeDim, int32_t cubeInc, const uint64_t *codes, int64_t ncodes,
                                const char *file, const char *dataset);
int vt_volume_write_hdf5(vt_context *ctx, const float *d_vol, int64_t X, int64_t Y, int64_t Z,
                         const char *file, const char *dataset, int32_t chunk);

/* ------------------------------------------- learned brick prefetcher */

/* Brick-visibility MLP (no reference counterpart): 12 features -> 256 ->
 * 256 -> 256 -> 1 logit, ReLU, BF16 weights and activations with FP32
 * accumulation on MFMA, FP32 master weights trained with Adam.
 * Parameters (vt_mlp_num_params floats, this order): W1[12][256], b1[256],
 * W2[256][256], b2[256], W3[256][256], b3[256], w4[256], b4[1]; layer
 * k computes h_k = relu(h_{k-1} W_k + b_k) and the logit is h_3 . w4 + b4. */
typedef struct vt_mlp vt_mlp;
int vt_mlp_num_params(void);
int vt_mlp_create(vt_context *ctx, uint64_t seed, int64_t max_batch, vt_mlp **out);
int vt_mlp_destroy(vt_mlp *m);
int vt_mlp_get_params(vt_mlp *m, float *out);
int vt_mlp_set_params(vt_mlp *m, const float *in);
/* logits of N feature rows d_x[N][12] (device) into d_logits[N] (device) */
int vt_mlp_forward(vt_mlp *m, const float *d_x, int64_t N, float *d_logits);
/* gradient of the mean weighted BCE-with-logits (labels d_y, positives
 * weighted by pos_weight) into host grad_out[vt_mlp_num_params()] */
int vt_mlp_gradients(vt_mlp *m, const float *d_x, const float *d_y, int64_t N, float pos_weight, float *grad_out,
                     float *loss_out);
/* one Adam step (beta 0.9/0.999, eps 1e-8) on the batch; loss_out may be NULL */
int vt_mlp_train_step(vt_mlp *m, const float *d_x, const float *d_y, int64_t N, float lr, float pos_weight,
                      float *loss_out);
/* features of (camera, brick) pairs: camera position and brick centre in
 * cube units minus 0.5, view direction, levelC/10, and the brick centre's
 * image-plane position (u, v) = 2*distance*(d.right / w, d.up / h) / (d.look)
 * with d = centre - camera (inside the frame for |u|, |v| <= 1), clamped to
 * [-4, 4], and 4 for a centre behind the camera. */
int vt_mlp_features(vt_mlp *m, const vt_camera *cam, int32_t nLevels, int32_t levelC, const uint64_t *d_codes,
                    int64_t N, float *d_x);
/* logits of the bricks d_codes[N] (device, level-levelC Morton codes) for the
 * camera, into d_logits[N] (device): the features above and the forward pass
 * fused into one launch (64 bricks per workgroup, activations in LDS), equal
 * bit for bit to vt_mlp_features followed by vt_mlp_forward.  N is not
 * bounded by max_batch.  vt_prefetch scores its candidates this way. */
int vt_mlp_score(vt_mlp *m, const vt_camera *cam, int32_t nLevels, int32_t levelC, const uint64_t *d_codes,
                 int64_t N, float *d_logits);

/* Score every candidate brick (the bricks holding octree nodes, i.e. the
 * level-levelC run list) for the camera with the MLP, then load the
 * highest-scoring non-resident ones with probability >= min_prob, at most
 * `budget`, into free slots or in place of resident bricks that score lower
 * and below probability 1 - min_prob.  The copies are issued on the
 * context's copy stream without waiting; the next vt_render_frame waits for
 * them before its first pass.
 * Hits are never affected: the render still decides every brick it marches.
 * *loaded (may be NULL) receives the number of bricks loaded. */
int vt_prefetch(vt_context *ctx, vt_mlp *m, const vt_camera *cam, int64_t budget, float min_prob, int64_t *loaded);

/* ------------------------------------------- frames in flight (serving) */

/* Frame requests served through the context's octree and out-of-core brick
 * cache with several frames in flight (no reference counterpart: the
 * reference renders one frame at a time, visualTur.cu:167-228).  A worker
 * thread takes submitted frames into up to max_frames slots; each launch
 * carries the parked rays of the frames already started (their later
 * passes) and the first pass of the frames admitted since the last one, and
 * the service between launches loads bricks for all of them.  With a
 * prefetcher, the next waiting frame's likely bricks are scored and copied
 * into free slots while the launch runs.  Each frame's pixels and hit
 * records equal vt_render_frame's for the same camera.  While a vt_frames
 * exists the context belongs to it: call no other vt_* function on that
 * context (the MLP given to vt_frames_set_prefetch included).  The march
 * options of the context at creation apply. */
typedef struct vt_frames vt_frames;
typedef struct {
    int64_t frames, launches;
    int64_t bricks_requested, bricks_uploaded, bricks_early, bricks_prefetched, host_tier_misses;
    double ms_kernel;           /* device time of the launches */
    double wall_s, busy_s;      /* since creation or the last vt_frames_get_stats; busy: frames in flight */
} vt_frames_stats;
/* max_frames: 1..32; max_rays: the largest W*H a frame may have. */
int vt_frames_create(vt_context *ctx, int32_t max_frames, int64_t max_rays, vt_frames **out);
/* m = NULL turns the prefetcher off; budget and min_prob as in vt_prefetch. */
int vt_frames_set_prefetch(vt_frames *f, vt_mlp *m, int64_t budget, float min_prob);
/* Queue one frame (outputs: caller-owned device memory, d_hits required,
 * untouched by the caller until the frame is done); every frame uses the
 * same levelO and levelC.  Returns a ticket. */
int vt_frames_submit(vt_frames *f, const vt_camera *cam, const vt_render_params *p, float *d_rgba,
                     vt_hit *d_hits, int64_t *ticket);
/* *done = 1 with the frame's status and statistics (passes: launches it took
 * part in; ms_total: submission to completion) and the ticket forgotten, or
 * *done = 0; wait = 1 blocks. */
int vt_frames_poll(vt_frames *f, int64_t ticket, int32_t wait, int32_t *done, vt_frame_stats *stats);
/* Totals since creation or the last call (which resets them). */
int vt_frames_get_stats(vt_frames *f, vt_frames_stats *out);
/* Serves what was submitted, then stops the worker. */
int vt_frames_destroy(vt_frames *f);

/* --------------------------------------------------------- serving loop */

/* Render-request serving (no reference counterpart).  Scenes keep every
 * brick resident in HBM plus a min/max pyramid over the octree levels, so a
 * request may carry its own isovalue (a node is valid for iso t exactly
 * when min < t <= max over its zero-padded samples).  A worker thread serves
 * requests as they are submitted (continuous batching): it admits a request
 * while its outputs (48 B per ray) and its isovalue's membership tables fit
 * one HBM arena of hbm_budget_bytes (tables no admitted request uses are
 * evicted, least recently used first), renders 16x16 tiles of all admitted
 * requests in one launch (a captured hipGraph, replayed per batch), and
 * keeps two batches in flight, so admission and the next batch's set-up
 * overlap the device work.  Mode 1 renders one request per launch in
 * arrival order (the plain per-request baseline).  The server has its own
 * HIP streams; scenes are added while it is stopped. */
typedef struct vt_server vt_server;
typedef struct {
    int32_t scene;          /* from vt_server_add_scene */
    vt_camera cam;
    float iso;
    float light[3];
    float color_scale;
    double arrival_s;       /* arrival time, seconds after the run starts */
    float *h_rgba;          /* optional host copy of the W*H float4 pixels */
    vt_hit *h_hits;         /* optional host copy of the hit records */
} vt_request;
typedef struct {
    double admit_s, done_s; /* seconds after the run starts */
    double latency_ms;      /* done - arrival */
    int32_t launches;       /* launches the request took part in */
    int32_t pad_;
} vt_request_result;
typedef struct {
    int64_t requests, launches, tiles, rays;
    double wall_s, busy_s;      /* busy: time with at least one batch on the device */
    int64_t tables_built;       /* isovalue membership tables built since the server was created */
    int64_t tables_held;        /* ... and held in the arena when the run stopped */
} vt_serve_stats;
int vt_server_create(vt_context *ctx, int64_t hbm_budget_bytes, int32_t max_tiles, vt_server **out);
int vt_server_destroy(vt_server *sv);
/* Lanes per ray of the serving launch (results are identical for every
 * setting): a batch gets 2^k lanes per ray (k <= 3), the largest with
 * rays x 2^k <= lane_budget (default 2097152), so light batches march each
 * ray with several lanes; 0 keeps one lane per ray; -k fixes 2^k. */
int vt_server_set_lanes(vt_server *sv, int64_t lane_budget);
/* A volume in HBM (C order [X][Y][Z]) becomes a scene: octree levels
 * 0..levelO, bricks of level levelC with cubeInc halo, all resident. */
int vt_server_add_scene(vt_server *sv, const float *d_vol, int32_t X, int32_t Y, int32_t Z, int32_t levelO,
                        int32_t levelC, int32_t cubeInc, int32_t *scene_out);
/* The same for requests with isovalues in [iso_lo, iso_hi] only: resident
 * are the bricks whose level-levelC min/max has min < iso_hi and
 * max >= iso_lo, which holds every brick such a request can march (a node
 * is visited only when min < iso <= max); requests outside the range are
 * refused (VT_E_ARG).  *resident_out (may be NULL) gets the number of
 * resident bricks. */
int vt_server_add_scene_range(vt_server *sv, const float *d_vol, int32_t X, int32_t Y, int32_t Z, int32_t levelO,
                              int32_t levelC, int32_t cubeInc, float iso_lo, float iso_hi, int32_t *scene_out,
                              int64_t *resident_out);
/* Online form.  vt_server_start runs the worker in `mode`; vt_server_submit
 * queues a request from any host thread (arrival_s is ignored: it arrives
 * now; h_rgba/h_hits must stay valid until it is done) and returns a ticket;
 * vt_server_poll reports it (wait=1 blocks): *done = 1 with the request's
 * status and times in seconds after its submission (the ticket is then
 * forgotten), or *done = 0.  vt_server_stop serves what was submitted, then
 * stops the worker and returns the run's statistics. */
int vt_server_start(vt_server *sv, int32_t mode);
int vt_server_submit(vt_server *sv, const vt_request *req, int64_t *ticket);
int vt_server_poll(vt_server *sv, int64_t ticket, int32_t wait, int32_t *done, vt_request_result *result);
int vt_server_stop(vt_server *sv, vt_serve_stats *stats);
/* Replay form: starts the worker, submits each request at its arrival time
 * (seconds after the call) from the calling thread, waits for all of them
 * and stops.  Result times are seconds after the call; a request that fails
 * (e.g. VT_E_RANGE: a sample outside its slot) has its status returned
 * after every request has finished, and the server stays usable. */
int vt_server_run(vt_server *sv, const vt_request *reqs, int64_t n, int32_t mode, vt_request_result *results,
                  vt_serve_stats *stats);

/* Streaming device copy (no reference counterpart): a kernel of known
 * traffic, read bytes + written bytes, that calibrates HBM byte counters. */
int vt_copy_probe(vt_context *ctx, const void *d_src, void *d_dst, size_t bytes);

/* Write a frame (host copy of the W*H float4 pixel buffer) as a binary PPM,
 * laid out as testVisualTur.cu:67-107 lays out its FreeImage bitmap: the
 * image is H pixels wide and W tall, pixel (x = i, y = j) holds ray
 * i*W + j, y counts from the bottom row, and each channel is the value
 * times 255 truncated to a byte (clamped to [0, 255]; NaN gives 0). */
int vt_image_write_ppm(const float *h_rgba, int32_t W, int32_t H, const char *path);

/* ---------------------------------------------------------------- files */

/* FileManager::FileManager (FileManager.cu:11-37), read-only. */
int vt_file_open(const char *file, const char *dataset, vt_file **out, int64_t dims[3]);
int vt_file_close(vt_file *f);
/* FileManager::readHDF5_Voxel_Array (FileManager.cu:86-167): box [s,e),
 * zero outside the dataset, C order, host memory. */
int vt_file_read_box(vt_file *f, const int32_t s[3], const int32_t e[3], float *out);
/* FileManager::CreateFile (FileManager.cu:169-243); chunk 0 = contiguous. */
int vt_file_create(const char *file, const char *dataset, const float *data,
                   int64_t X, int64_t Y, int64_t Z, int32_t chunk);
/* The cutFile tool (cutFile.cu:13-21): the box [x, x+dim) x [y, y+dim) x
 * [z, z+dim) of `dataset` (zero outside it, as readHDF5_Voxel_Array reads)
 * written to `out_file` as a contiguous dim^3 dataset named "value". */
int vt_file_cut(const char *file, const char *dataset, const char *out_file, int32_t x, int32_t y, int32_t z,
                int32_t dim);

#ifdef __cplusplus
}
#endif
#endif
