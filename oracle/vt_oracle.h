/*
 * vt_oracle.h -- CPU oracle for the visualTur isosurface hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker (or as the timed CPU
 * baseline), never as the thing measured or shipped.
 *
 * It restates, in plain C, the reference algorithm of carlosduelo/visualTur:
 *   Morton codes            inc/mortonCodeUtil.hpp:21-181
 *   run-list membership     src/Octree.cu:15-142, 375-532; src/CreateOctree.cu:130-284
 *   slab ray/AABB test      src/Octree.cu:258-330 (== src/rayCaster.cu:12-79)
 *   stack DFS traversal     src/Octree.cu:620-760
 *   trilinear march/secant  src/rayCaster.cu:81-237
 *   normal + diffuse shade  src/rayCaster.cu:131-137, 238-251
 *   pixel ray generation    src/Camera.cu:58-73, src/Screen.cpp:11-24
 *   octree builder          src/CreateOctree.cu:79-128, 300-371, 425-573
 *   octree file format      src/Octree.cu:775-815, src/CreateOctree.cu:342-371
 *   zero-padded box read    src/FileManager.cu:86-167
 *
 * Floating point: compiled with -ffp-contract=off, no fast-math, so every
 * float expression below rounds exactly as written.  The normalize() of the
 * reference uses CUDA rsqrtf; this oracle (and the HIP kernels it checks)
 * use 1.0f/sqrtf(x), both correctly rounded.
 */
#ifndef VT_ORACLE_H
#define VT_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- Morton */
uint64_t vto_dilate(uint64_t x);
uint64_t vto_encode(int x, int y, int z, int level, int nLevels);
int      vto_level(uint64_t code);
void     vto_decode(uint64_t code, int level, int nLevels, int32_t out[3]);
/* Incremental min-box update, mortonCodeUtil.hpp:151-181. */
void     vto_update_coordinates(int maxLevel, int cLevel, uint64_t cIndex,
                                int nLevel, uint64_t nIndex,
                                const int32_t minBox[3], int32_t out[3]);

/* ------------------------------------------------------ run-list search */
/* Membership mask (bit k = child (parent<<3)+k) as the traversal kernel
 * computes it, Octree.cu:375-532 (windowed binary search + sequential). */
uint32_t vto_children_mask_device(const uint64_t *runs, int size, uint64_t parent);
/* Same, as the offline builder computes it, CreateOctree.cu:182-284. */
uint32_t vto_children_mask_builder(const uint64_t *runs, int size, uint64_t parent);
/* Plain set membership over sorted inclusive run pairs (checker of the two
 * restatements above). */
int      vto_member_exact(const uint64_t *runs, int size, uint64_t code);

/* ------------------------------------------------------------------ slab */
int vto_slab(const float o[3], const float d[3], const int32_t mn[3],
             const int32_t mx[3], float *tnear, float *tfar);

/* ---------------------------------------------------------------- camera */
typedef struct {
    float pos[3];
    float look[3];
    float up[3];
    float right[3];
    float h, w, distance;   /* screen plane size in world units, Screen.cpp:18-19 */
    int32_t W, H;
} vto_camera;

/* Screen::Screen plane size, Screen.cpp:18-19 (fov is a half-angle). */
void vto_screen_size(float distance, float fovH, float fovW, float *h, float *w);
/* Camera.cu:83-156: one camera move.  op: 0 Move(arg[0..2]), 1 RotateX,
 * 2 RotateY, 3 RotateZ (degrees, arg[0]), 4 MoveForward, 5 MoveUpward,
 * 6 StrafeRight (distance, arg[0]). */
void vto_camera_move(vto_camera *c, int op, const float *arg);
void vto_ray_dir(const vto_camera *cam, int i, int j, float out[3]);

/* ---------------------------------------------------------------- volume */
typedef struct {
    const float *data;   /* C order [X][Y][Z], z fastest */
    int32_t dims[3];
    int32_t origin[3];   /* voxel (x,y,z) is data[x - origin[0]][...]; 0 for a whole volume */
} vto_volume;

float vto_voxel(const vto_volume *v, int x, int y, int z);
float vto_sample(const vto_volume *v, float px, float py, float pz);
/* FileManager::readHDF5_Voxel_Array semantics on an in-memory volume:
 * out is (e-s) boxed, C order, zero outside the dataset. */
void  vto_read_box(const vto_volume *v, const int32_t s[3], const int32_t e[3], float *out);

/* ---------------------------------------------------------------- octree */
typedef struct {
    int32_t nLevels;        /* n, cube side D = 2^n */
    int32_t maxLevel;       /* levels 0..maxLevel present in runs/sizes */
    const uint64_t *runs;   /* levels 0..maxLevel concatenated */
    const int32_t *sizes;   /* maxLevel+1 entries */
} vto_octree;

/* Reference DFS (explicit stack).  Writes up to max_out emitted nodes of
 * level finalLevel, in emission order.  Returns how many were emitted in
 * total if the ray were run to exhaustion, capped at cap_total. */
int vto_traverse(const vto_octree *oct, const float o[3], const float d[3],
                 int finalLevel, uint64_t *out, int max_out, int cap_total);

/* ----------------------------------------------------------------- frame */
typedef struct {
    int32_t levelO;        /* octree level the traversal emits */
    int32_t levelC;        /* brick level of the cache */
    float iso;
    float light[3];        /* (0,512,0), visualTur.cu:30 */
    float color_scale;     /* 256.0f, rayCaster.cu:245 */
} vto_params;

typedef struct {
    uint64_t node;         /* levelO Morton code of the node hit, 0 on miss */
    uint64_t brick;        /* levelC brick code, 0 on miss */
    int32_t  voxel[3];     /* floor(Xnew) */
    float    depth;        /* |Xnew - origin| */
} vto_hit;

/* Render rays with pixel ids [ray0, ray1) of the frame.  rgba and hits are
 * indexed by pixel id (caller passes full-frame arrays).  If brick_seen is
 * non-null it must hold 8^levelC bytes; entry (brick - 8^levelC) is set to 1
 * for every brick marched by at least one ray (the visited-brick trace).
 * Uses up to nthreads OpenMP threads. */
int vto_render(const vto_volume *vol, const vto_octree *oct,
               const vto_camera *cam, const vto_params *p,
               int64_t ray0, int64_t ray1,
               float *rgba, vto_hit *hits, uint8_t *brick_seen, int nthreads);

/* Same, for an explicit list of pixel ids (sampled checks). */
int vto_render_list(const vto_volume *vol, const vto_octree *oct,
                    const vto_camera *cam, const vto_params *p,
                    const int64_t *ids, int64_t count,
                    float *rgba, vto_hit *hits, int nthreads);

/* A brick source: the zero-padded box [s, e) into out (C order); 0 = ok. */
typedef int (*vto_box_reader)(void *user, const int32_t s[3], const int32_t e[3], float *out);
int vto_volume_box_reader(void *user, const int32_t s[3], const int32_t e[3], float *out);

typedef struct {
    int64_t passes, hits, misses, evictions;
} vto_cache_stats;

/* FileManager (FileManager.cu:11-37, 86-167) over libhdf5, as a brick source
 * (vt_oracle_h5.c); NULL when the file or dataset cannot be opened. */
typedef struct vto_h5 vto_h5;
vto_h5 *vto_h5_open(const char *file, const char *dataset);
void vto_h5_close(vto_h5 *h);
int vto_h5_box_reader(void *user, const int32_t s[3], const int32_t e[3], float *out);

/* An LRU brick cache of `capacity` bricks of level level_c (cube^3 cells
 * plus an `inc` halo), kept across frames (lruCache.cu:101-129). */
typedef struct vto_cache vto_cache;
vto_cache *vto_cache_new(int level_c, int cube, int inc, int64_t capacity);
void vto_cache_free(vto_cache *c);

/* The reference frame loop (visualTur.cu:167-228) through cache `c`, filled
 * from `read` (lruCache.cu:168-313); the march samples the cached slots.
 * Same pixels and hit records as vto_render.  0 ok, -1 bad arguments,
 * -2 no memory, -3 the source failed. */
int vto_render_cached(const vto_octree *oct, const vto_camera *cam, const vto_params *p, vto_cache *c,
                      vto_box_reader read, void *user, float *rgba, vto_hit *hits, vto_cache_stats *stats,
                      int nthreads);

/* rayCaster::renderCube (rayCaster.cu:267-357): one brick, every pixel. */
int vto_render_cube(const vto_volume *vol, const vto_camera *cam, float iso, const float light[3],
                    const int32_t minBox[3], int dim, float *rgba);

/* --------------------------------------------------------------- builder */
/* CreateOctree restatement.  literal=1 reproduces the inner-level loop that
 * skips the last node of each level (CreateOctree.cu:550); literal=0 visits
 * it.  Returns the flattened run arrays for levels 0..nLevels through
 * malloc'ed *runs_out (free with vto_free) and sizes_out[nLevels+1].
 * Returns nLevels, or a negative value on error. */
int  vto_build_octree(const vto_volume *vol, float iso, int literal,
                      uint64_t **runs_out, int32_t *sizes_out, int max_levels);
void vto_free(void *p);

/* Octree file I/O (reference binary layout). */
int vto_octree_write(const char *path, float iso, int32_t dimension,
                     const int32_t realDim[3], int32_t nLevels,
                     const uint64_t *runs, const int32_t *sizes);
/* Reads header into iso/dimension/realDim/nLevels, and, if runs_out is
 * non-null, the flattened runs of levels 0..nLevels (malloc'ed). */
int  vto_octree_read(const char *path, float *iso, int32_t *dimension,
                     int32_t realDim[3], int32_t *nLevels,
                     uint64_t **runs_out, int32_t *sizes_out, int max_levels);

#ifdef __cplusplus
}
#endif
#endif
