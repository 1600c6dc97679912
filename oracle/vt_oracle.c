/*
 * vt_oracle.c -- CPU oracle for the visualTur isosurface hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see vt_oracle.h).  It restates the reference
 * algorithm; each function names the reference lines it follows.  It is
 * never linked into the shipped library.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fopenmp -fPIC -shared
 */
#define _GNU_SOURCE
#include "vt_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define VTO_MAGIC 919278872   /* Octree.cu:781, CreateOctree.cu:347 */

/* float -> int conversion with the CUDA/AMDGPU hardware semantics
 * (truncate, saturate, NaN -> 0).  C leaves NaN/out-of-range undefined. */
static int cvt_rzi(double v)
{
    if (v != v) return 0;
    if (v >= 2147483647.0) return INT_MAX;
    if (v <= -2147483648.0) return INT_MIN;
    return (int)v;
}

/* ================================================================ Morton */

/* mortonCodeUtil.hpp:10-19: spread the 21 low bits of x to every third bit.
 * Written from the definition (bit b goes to bit 3b). */
uint64_t vto_dilate(uint64_t x)
{
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b)
        r |= ((x >> b) & 1ull) << (3 * b);
    return r;
}

/* mortonCodeUtil.hpp:125-148 (coordinateToIndex). */
uint64_t vto_encode(int x, int y, int z, int level, int nLevels)
{
    if (level == 0) return 1;
    uint64_t code = 1ull << (3 * nLevels);
    code |= vto_dilate((uint64_t)x) << 2;
    code |= vto_dilate((uint64_t)y) << 1;
    code |= vto_dilate((uint64_t)z);
    return code >> (3 * (nLevels - level));
}

/* mortonCodeUtil.hpp:21-65 (getIndexLevel): largest l <= 19 with code >= 8^l,
 * 0 for code == 1, -1 otherwise. */
int vto_level(uint64_t code)
{
    for (int l = 19; l >= 1; --l)
        if (code >= (1ull << (3 * l))) return l;
    return code == 1 ? 0 : -1;
}

/* mortonCodeUtil.hpp:99-123 (getMinBoxIndex2). */
void vto_decode(uint64_t code, int level, int nLevels, int32_t out[3])
{
    out[0] = out[1] = out[2] = 0;
    if (code == 1) return;
    for (int l = level; l > 0; --l) {
        out[2] += (int32_t)((code & 1ull) << (nLevels - l)); code >>= 1;
        out[1] += (int32_t)((code & 1ull) << (nLevels - l)); code >>= 1;
        out[0] += (int32_t)((code & 1ull) << (nLevels - l)); code >>= 1;
    }
}

/* mortonCodeUtil.hpp:151-181 (updateCoordinates). */
void vto_update_coordinates(int maxLevel, int cLevel, uint64_t cIndex,
                            int nLevel, uint64_t nIndex,
                            const int32_t minBox[3], int32_t out[3])
{
    int32_t m[3] = {minBox[0], minBox[1], minBox[2]};
    if (nIndex == 0) { out[0] = out[1] = out[2] = 0; return; }
    if (cLevel > nLevel) { vto_decode(nIndex, nLevel, maxLevel, out); return; }
    m[2] += (int32_t)((nIndex & 1ull) << (maxLevel - nLevel)); nIndex >>= 1;
    m[1] += (int32_t)((nIndex & 1ull) << (maxLevel - nLevel)); nIndex >>= 1;
    m[0] += (int32_t)((nIndex & 1ull) << (maxLevel - nLevel));
    if (cLevel == nLevel) {
        m[2] -= (int32_t)((cIndex & 1ull) << (maxLevel - cLevel)); cIndex >>= 1;
        m[1] -= (int32_t)((cIndex & 1ull) << (maxLevel - cLevel)); cIndex >>= 1;
        m[0] -= (int32_t)((cIndex & 1ull) << (maxLevel - cLevel));
    }
    out[0] = m[0]; out[1] = m[1]; out[2] = m[2];
}

/* ===================================================== run-list searches */

static int guard_index(int i, int size)
{
    if (i < 0 || i >= size) {
        fprintf(stderr, "vt_oracle: run-list index %d out of [0,%d)\n", i, size);
        abort();
    }
    return i;
}

/* Octree.cu:15-20 / CreateOctree.cu:130-139 (inclusive pair test). */
static int check_range(const uint64_t *E, int size, uint64_t idx, int a, int b)
{
    uint64_t lo = E[guard_index(a, size)], hi = E[guard_index(b, size)];
    return idx == lo || idx == hi || (lo < idx && hi > idx);
}

/* Octree.cu:78-99 (_cuda_binary_search_closer, the active #if 1 branch). */
static int closer_device(const uint64_t *E, int size, uint64_t idx, int lo, int hi)
{
    int end = 0, found = 0, middle = 0;
    while (!end && !found) {
        int diff = hi - lo;
        middle = lo + diff / 2;
        if (middle % 2 == 1) middle--;
        end = diff <= 1;
        found = check_range(E, size, idx, middle, middle + 1);
        if (idx < E[guard_index(middle, size)]) hi = middle - 1;
        else lo = middle + 2;
    }
    return middle;
}

/* Octree.cu:134-142 (_cuda_searchSecuential). */
static int sequential_device(const uint64_t *E, int size, uint64_t idx, int lo, int hi)
{
    int find = 0;
    for (int i = lo; i < hi; i += 2)
        if (check_range(E, size, idx, i, i + 1)) find = 1;
    return find;
}

/* Membership part of Octree.cu:375-532 (_cuda_searchChildrenValidAndHit2):
 * size==2 uses the single pair; otherwise both window ends are found by
 * closer-search from the first child and child+7, then scanned.  A level
 * with no runs (size 0) has no members. */
uint32_t vto_children_mask_device(const uint64_t *E, int size, uint64_t parent)
{
    uint64_t first = parent << 3;
    uint32_t mask = 0;
    if (size <= 0) return 0;
    if (size == 2) {
        for (int k = 0; k < 8; ++k)
            if (check_range(E, size, first + (uint64_t)k, 0, 1)) mask |= 1u << k;
        return mask;
    }
    unsigned c1 = (unsigned)closer_device(E, size, first, 0, size - 1);
    unsigned c8 = (unsigned)closer_device(E, size, first + 7, (int)c1, size - 1) + 1;
    if (c8 >= (unsigned)size) c8 = (unsigned)size - 1;
    for (int k = 0; k < 8; ++k)
        if (sequential_device(E, size, first + (uint64_t)k, (int)c1, (int)c8)) mask |= 1u << k;
    return mask;
}

/* CreateOctree.cu:182-221 (OctreeLevel::binary_search_closer, recursive). */
static int closer_builder(const uint64_t *E, int size, uint64_t idx, int lo, int hi)
{
    for (;;) {
        int diff = hi - lo;
        unsigned middle = (unsigned)(lo + (diff / 2));
        if (diff <= 1) {
            if (middle % 2 == 1) middle--;
            return (int)middle;
        }
        if (middle % 2 == 1) middle--;
        if (check_range(E, size, idx, (int)middle, (int)middle + 1)) return (int)middle;
        if (idx < E[guard_index((int)middle, size)]) hi = (int)middle - 1;
        else if (idx > E[guard_index((int)middle + 1, size)]) lo = (int)middle + 2;
        else return 0;   /* unreachable "Errro" branch, returns false */
    }
}

/* CreateOctree.cu:223-244 (searchSecuential, up=true, with early exit). */
static int sequential_builder(const uint64_t *E, int size, uint64_t idx, int lo, int hi)
{
    for (int i = lo; i < hi; i += 2) {
        if (check_range(E, size, idx, i, i + 1)) return 1;
        if (idx < E[guard_index(i + 1, size)]) break;
    }
    return 0;
}

/* CreateOctree.cu:247-284 (OctreeLevel::searchChildren). */
uint32_t vto_children_mask_builder(const uint64_t *E, int size, uint64_t parent)
{
    uint64_t first = parent << 3;
    uint32_t mask = 0;
    if (size <= 0) return 0;
    if (size == 2) {
        for (int k = 0; k < 8; ++k)
            if (check_range(E, size, first + (uint64_t)k, 0, 1)) mask |= 1u << k;
        return mask;
    }
    unsigned c1 = (unsigned)closer_builder(E, size, first, 0, size);
    unsigned c8 = (unsigned)closer_builder(E, size, first + 7, (int)c1, size) + 1;
    if (c8 >= (unsigned)size) c8 = (unsigned)size - 1;
    for (int k = 0; k < 8; ++k)
        if (sequential_builder(E, size, first + (uint64_t)k, (int)c1, (int)c8)) mask |= 1u << k;
    return mask;
}

int vto_member_exact(const uint64_t *E, int size, uint64_t code)
{
    int lo = 0, hi = size / 2 - 1, best = -1;
    while (lo <= hi) {              /* last pair whose start <= code */
        int mid = (lo + hi) / 2;
        if (E[2 * mid] <= code) { best = mid; lo = mid + 1; }
        else hi = mid - 1;
    }
    return best >= 0 && code <= E[2 * best + 1];
}

/* =================================================================== slab */

/* Octree.cu:258-330 (_cuda_RayAABB2); identical arithmetic in
 * Octree.cu:182-256 and rayCaster.cu:12-79. */
int vto_slab(const float o[3], const float d[3], const int32_t mn[3],
             const int32_t mx[3], float *tnear, float *tfar)
{
    int hit = 1;
    float tmin, tmax, tymin, tymax, tzmin, tzmax;
    float divx = 1 / d[0];
    if (divx >= 0) { tmin = ((float)mn[0] - o[0]) * divx; tmax = ((float)mx[0] - o[0]) * divx; }
    else           { tmin = ((float)mx[0] - o[0]) * divx; tmax = ((float)mn[0] - o[0]) * divx; }
    float divy = 1 / d[1];
    if (divy >= 0) { tymin = ((float)mn[1] - o[1]) * divy; tymax = ((float)mx[1] - o[1]) * divy; }
    else           { tymin = ((float)mx[1] - o[1]) * divy; tymax = ((float)mn[1] - o[1]) * divy; }
    if ((tmin > tymax) || (tymin > tmax)) hit = 0;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float divz = 1 / d[2];
    if (divz >= 0) { tzmin = ((float)mn[2] - o[2]) * divz; tzmax = ((float)mx[2] - o[2]) * divz; }
    else           { tzmin = ((float)mx[2] - o[2]) * divz; tzmax = ((float)mn[2] - o[2]) * divz; }
    if ((tmin > tzmax) || (tzmin > tmax)) hit = 0;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    *tnear = (tmin < 0.0) ? 0.0f : tmin;
    *tfar = tmax;
    return hit;
}

/* ================================================================= camera */

/* Screen.cpp:18-19: the fov argument is used as a half-angle. */
void vto_screen_size(float distance, float fovH, float fovW, float *h, float *w)
{
    *h = 2 * distance * tanf((float)(fovH * (M_PI / 180.0)));
    *w = 2 * distance * tanf((float)(fovW * (M_PI / 180.0)));
}

/* float3 arithmetic of cutil_math, component by component. */
static void v3_axpby(const float a[3], float s, const float b[3], float t, float sign, float out[3])
{
    /* out = a*s + sign*(b*t), rounded as the reference's a*s + b*t / a*s - b*t */
    for (int k = 0; k < 3; ++k) {
        float x = a[k] * s, y = b[k] * t;
        out[k] = sign > 0 ? x + y : x - y;
    }
}

static void v3_cross(const float a[3], const float b[3], float out[3])
{
    float r0 = a[1] * b[2] - a[2] * b[1];
    float r1 = a[2] * b[0] - a[0] * b[2];
    float r2 = a[0] * b[1] - a[1] * b[0];
    out[0] = r0; out[1] = r1; out[2] = r2;
}

/* Camera.cu:83-156.  The angle's sine and cosine are taken in double and
 * stored as float (:91-92).  RotateX/Y/Z call normalize() without using its
 * result (:97, :113, :129), so the rotated vector is not renormalised.
 * MoveUpward moves along `right` and StrafeRight along `up` (:146-156). */
void vto_camera_move(vto_camera *c, int op, const float *arg)
{
    float t[3];
    if (op == 0) {                 /* Move: position += Direction */
        for (int k = 0; k < 3; ++k) c->pos[k] = c->pos[k] + arg[k];
        return;
    }
    if (op >= 1 && op <= 3) {
        float s = (float)sin(arg[0] * (M_PI / 180.0));
        float co = (float)cos(arg[0] * (M_PI / 180.0));
        if (op == 1) {             /* look = look*c + up*s; up = -cross(look, right) */
            v3_axpby(c->look, co, c->up, s, 1.0f, t);
            memcpy(c->look, t, sizeof t);
            v3_cross(c->look, c->right, t);
            for (int k = 0; k < 3; ++k) c->up[k] = (-1.0f) * t[k];
        } else if (op == 2) {      /* look = look*c - right*s; right = cross(look, up) */
            v3_axpby(c->look, co, c->right, s, -1.0f, t);
            memcpy(c->look, t, sizeof t);
            v3_cross(c->look, c->up, c->right);
        } else {                   /* right = right*c + up*s; up = -cross(look, right) */
            v3_axpby(c->right, co, c->up, s, 1.0f, t);
            memcpy(c->right, t, sizeof t);
            v3_cross(c->look, c->right, t);
            for (int k = 0; k < 3; ++k) c->up[k] = (-1.0f) * t[k];
        }
        return;
    }
    const float *dir = op == 4 ? c->look : op == 5 ? c->right : c->up;
    const float d = op == 4 ? -arg[0] : arg[0];
    for (int k = 0; k < 3; ++k) c->pos[k] = c->pos[k] + dir[k] * d;
}

/* Camera.cu:58-73 (the active #else branch): row i counts from the bottom,
 * pixel id = i*W + j.  normalize() as v * (1/sqrt(dot(v,v))). */
void vto_ray_dir(const vto_camera *c, int i, int j, float out[3])
{
    float ih = c->h / (float)c->H;
    float iw = c->w / (float)c->W;
    float su = -(c->h / 2.0f) + (ih * ((float)i + 0.5f));
    float sr = -(c->w / 2.0f) + (iw * ((float)j + 0.5f));
    float A[3];
    for (int k = 0; k < 3; ++k) A[k] = c->look[k] * c->distance;
    for (int k = 0; k < 3; ++k) A[k] = A[k] + c->up[k] * su;
    for (int k = 0; k < 3; ++k) A[k] = A[k] + c->right[k] * sr;
    float len2 = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
    float inv = 1.0f / sqrtf(len2);
    for (int k = 0; k < 3; ++k) out[k] = A[k] * inv;
}

/* ================================================================= volume */

/* Zero-padded sample at an integer point (FileManager.cu:88-112 fills zeros
 * outside the dataset before the hyperslab read). */
float vto_voxel(const vto_volume *v, int x, int y, int z)
{
    x -= v->origin[0];
    y -= v->origin[1];
    z -= v->origin[2];
    if (x < 0 || y < 0 || z < 0 || x >= v->dims[0] || y >= v->dims[1] || z >= v->dims[2])
        return 0.0f;
    return v->data[((size_t)x * (size_t)v->dims[1] + (size_t)y) * (size_t)v->dims[2] + (size_t)z];
}

static int coord_ok(float p)
{
    return fabsf(p) < 16777216.0f;   /* finite and exactly integral-convertible */
}

/* rayCaster.cu:86-106 (getElementInterpolate), reading the global
 * zero-padded volume instead of a brick slot; the slot with halo holds the
 * same values wherever the reference reads inside it.  A non-finite or huge
 * coordinate (the reference would read outside the slot) yields NaN. */
float vto_sample(const vto_volume *v, float px, float py, float pz)
{
    if (!coord_ok(px) || !coord_ok(py) || !coord_ok(pz)) return NAN;
    float ix, iy, iz;
    float fx = modff(px, &ix);
    float fy = modff(py, &iy);
    float fz = modff(pz, &iz);
    int x0 = (int)ix, y0 = (int)iy, z0 = (int)iz;
    int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    float c00 = vto_voxel(v, x0, y0, z0) * (1.0f - fx) + vto_voxel(v, x1, y0, z0) * fx;
    float c01 = vto_voxel(v, x0, y0, z1) * (1.0f - fx) + vto_voxel(v, x1, y0, z1) * fx;
    float c10 = vto_voxel(v, x0, y1, z0) * (1.0f - fx) + vto_voxel(v, x1, y1, z0) * fx;
    float c11 = vto_voxel(v, x0, y1, z1) * (1.0f - fx) + vto_voxel(v, x1, y1, z1) * fx;
    float c0 = c00 * (1.0f - fy) + c10 * fy;
    float c1 = c01 * (1.0f - fy) + c11 * fy;
    return c0 * (1.0f - fz) + c1 * fz;
}

/* FileManager.cu:86-167 on an in-memory dataset. */
void vto_read_box(const vto_volume *v, const int32_t s[3], const int32_t e[3], float *out)
{
    int dx = abs(e[0] - s[0]), dy = abs(e[1] - s[1]), dz = abs(e[2] - s[2]);
    memset(out, 0, (size_t)dx * dy * dz * sizeof(float));
    if (s[0] >= v->dims[0] || s[1] >= v->dims[1] || s[2] >= v->dims[2] ||
        e[0] < 0 || e[1] < 0 || e[2] < 0)
        return;
    for (int x = 0; x < dx; ++x)
        for (int y = 0; y < dy; ++y)
            for (int z = 0; z < dz; ++z)
                out[((size_t)x * dy + y) * dz + z] = vto_voxel(v, s[0] + x, s[1] + y, s[2] + z);
}

/* ============================================================= traversal */

#define VTO_STACK 256   /* the reference uses 32 (Octree.cu:7) without checks */

typedef struct {
    const vto_octree *oct;
    const uint64_t *level_runs[33];
    float o[3], d[3];
} trav_ctx;

static void trav_init(trav_ctx *t, const vto_octree *oct, const float o[3], const float d[3])
{
    t->oct = oct;
    size_t off = 0;
    for (int l = 0; l <= oct->maxLevel && l < 33; ++l) {
        t->level_runs[l] = oct->runs + off;
        off += (size_t)oct->sizes[l];
    }
    memcpy(t->o, o, sizeof t->o);
    memcpy(t->d, d, sizeof t->d);
}

/* Root acceptance, Octree.cu:729: member of level 0 and slab-hit (no tfar test). */
static int root_ok(const trav_ctx *t)
{
    const vto_octree *oct = t->oct;
    if (oct->sizes[0] < 2) return 0;
    if (!check_range(t->level_runs[0], oct->sizes[0], 1, 0, 1)) return 0;
    int32_t mn[3] = {0, 0, 0}, side = 1 << oct->nLevels;
    int32_t mx[3] = {side, side, side};
    float tn, tf;
    return vto_slab(t->o, t->d, mn, mx, &tn, &tf);
}

/* Octree.cu:620-720 (_cuda_getFirtsVoxel) over an explicit stack.
 * Returns 1 and the emitted node, or 0 when the stack is exhausted. */
typedef struct {
    uint64_t idx[VTO_STACK];
    int lvl[VTO_STACK];
    int top;
} trav_stack;

/* The DFS over a stack of `cap` entries (sidx, slvl, *stop). */
static int trav_next_s(const trav_ctx *t, uint64_t *sidx, int *slvl, int *stop, int cap, int finalLevel,
                       uint64_t *out)
{
    const vto_octree *oct = t->oct;
    int n = oct->nLevels;
    while (*stop >= 0) {
        if (slvl[*stop] == finalLevel) {
            *out = sidx[*stop];
            (*stop)--;
            return 1;
        }
        uint64_t current = sidx[*stop];
        int clevel = slvl[*stop];
        (*stop)--;
        int32_t minBox[3];
        vto_decode(current, clevel, n, minBox);
        int level = clevel + 1;
        uint32_t members = (level <= oct->maxLevel)
            ? vto_children_mask_device(t->level_runs[level], oct->sizes[level], current) : 0;
        int dim = 1 << (n - level);
        uint64_t children[8];
        float tnears[8], tfars[8];
        int nValid = 0;
        for (int k = 0; k < 8; ++k) {
            int32_t mn[3] = {minBox[0] + ((k >> 2) & 1) * dim,
                             minBox[1] + ((k >> 1) & 1) * dim,
                             minBox[2] + (k & 1) * dim};
            int32_t mx[3] = {mn[0] + dim, mn[1] + dim, mn[2] + dim};
            float tn, tf;
            int hit = vto_slab(t->o, t->d, mn, mx, &tn, &tf);
            if (hit && tf >= 0.0 && ((members >> k) & 1)) {
                children[nValid] = (current << 3) + (uint64_t)k;
                tnears[nValid] = tn;
                tfars[nValid] = tf;
                nValid++;
            }
        }
        /* Octree.cu:665-706: bubble sort by tnear, ties by tfar. */
        int nn = nValid;
        while (nn != 0) {
            int newn = 0;
            for (int id = 1; id < nValid; ++id) {
                int sw = (tnears[id - 1] > tnears[id]) ||
                         (tnears[id - 1] == tnears[id] && tfars[id - 1] > tfars[id]);
                if (sw) {
                    uint64_t a = children[id]; children[id] = children[id - 1]; children[id - 1] = a;
                    float f = tnears[id]; tnears[id] = tnears[id - 1]; tnears[id - 1] = f;
                    f = tfars[id]; tfars[id] = tfars[id - 1]; tfars[id - 1] = f;
                    newn = id;
                }
            }
            nn = newn;
        }
        /* Octree.cu:707-712: push in reverse so the nearest is on top. */
        for (int i = nValid - 1; i >= 0; --i) {
            if (*stop + 1 >= cap) { fprintf(stderr, "vt_oracle: stack overflow\n"); abort(); }
            (*stop)++;
            sidx[*stop] = children[i];
            slvl[*stop] = level;
        }
    }
    return 0;
}

static int trav_next(const trav_ctx *t, trav_stack *st, int finalLevel, uint64_t *out)
{
    return trav_next_s(t, st->idx, st->lvl, &st->top, VTO_STACK, finalLevel, out);
}

static void trav_reset(trav_stack *st)
{
    /* Octree.cu:750-760 (cuda_resetState). */
    st->top = 0;
    st->idx[0] = 1;
    st->lvl[0] = 0;
}

int vto_traverse(const vto_octree *oct, const float o[3], const float d[3],
                 int finalLevel, uint64_t *out, int max_out, int cap_total)
{
    trav_ctx t;
    trav_init(&t, oct, o, d);
    if (!root_ok(&t)) return 0;
    trav_stack *st = (trav_stack *)malloc(sizeof(trav_stack));
    trav_reset(st);
    int count = 0;
    uint64_t node;
    while (count < cap_total && trav_next(&t, st, finalLevel, &node)) {
        if (count < max_out) out[count] = node;
        count++;
    }
    free(st);
    return count;
}

/* ================================================================= march */

/* rayCaster.cu:156-237 for one node: returns 1 on a crossing with the
 * secant point in Xnew.  The step count cap (4*side+8) lies above the
 * geometric bound 2*side*sqrt(3)+1 and only bounds degenerate inputs. */
static int march_node(const vto_volume *v, const float o[3], const float d[3],
                      const int32_t mn[3], int side, float iso, float Xnew[3])
{
    int32_t mx[3] = {mn[0] + side, mn[1] + side, mn[2] + side};
    float tnear, tfar;
    if (!vto_slab(o, d, mn, mx, &tnear, &tfar)) return 0;
    float X[3], Xfar[3], vStep[3];
    for (int k = 0; k < 3; ++k) X[k] = o[k] + tnear * d[k];
    for (int k = 0; k < 3; ++k) { Xfar[k] = X[k]; Xnew[k] = X[k]; }
    for (int k = 0; k < 3; ++k) vStep[k] = 0.5f * d[k];
    int maxStep = cvt_rzi(ceil((double)(tfar - tnear) / 0.5));
    if (maxStep > 4 * side + 8) maxStep = 4 * side + 8;
    int first = 1, hit = 0;
    float ant = 0.0f, sig = 0.0f;
    int steps = 0;
    while (steps <= maxStep) {
        if (first) {
            first = 0;
            ant = vto_sample(v, X[0], X[1], X[2]);
            for (int k = 0; k < 3; ++k) Xfar[k] = X[k];
        } else {
            sig = vto_sample(v, X[0], X[1], X[2]);
            if (((iso - ant) < 0.0 && (iso - sig) < 0.0) || ((iso - ant) > 0.0 && (iso - sig) > 0.0)) {
                ant = sig;
                for (int k = 0; k < 3; ++k) Xfar[k] = X[k];
            } else {
                float a = (iso - ant) / (sig - ant);
                for (int k = 0; k < 3; ++k) Xnew[k] = Xfar[k] * (1.0f - a) + X[k] * a;
                hit = 1;
                steps = maxStep;
            }
        }
        for (int k = 0; k < 3; ++k) X[k] = X[k] + vStep[k];
        steps++;
    }
    return hit;
}

static void normalize3(float v[3])
{
    float inv = 1.0f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] = v[0] * inv; v[1] = v[1] * inv; v[2] = v[2] * inv;
}

/* rayCaster.cu:131-137 (getNormal) and 238-251 (diffuse shade). */
static void shade(const vto_volume *v, const float Xn[3], const vto_params *p, float rgba[4])
{
    float n[3];
    n[0] = (vto_sample(v, Xn[0] - 1.0f, Xn[1], Xn[2]) - vto_sample(v, Xn[0] + 1.0f, Xn[1], Xn[2])) / 2.0f;
    n[1] = (vto_sample(v, Xn[0], Xn[1] - 1.0f, Xn[2]) - vto_sample(v, Xn[0], Xn[1] + 1.0f, Xn[2])) / 2.0f;
    n[2] = (vto_sample(v, Xn[0], Xn[1], Xn[2] - 1.0f) - vto_sample(v, Xn[0], Xn[1], Xn[2] + 1.0f)) / 2.0f;
    normalize3(n);
    float l[3] = {Xn[0] - p->light[0], Xn[1] - p->light[1], Xn[2] - p->light[2]};
    normalize3(l);
    float dif = fabsf(n[0] * l[0] + n[1] * l[1] + n[2] * l[2]);
    float a = Xn[1] / p->color_scale;
    rgba[0] = (1 - a) * dif;
    rgba[1] = (a) * dif;
    rgba[2] = 0.0f * dif;
    rgba[3] = 1.0f;
}

static int floor_to_int(float x)
{
    return cvt_rzi((double)floorf(x));
}

/* One ray of the frame loop (visualTur.cu:167-228 with the pass state
 * machine collapsed: the result of a ray does not depend on where the pass
 * boundaries fall when the brick cache returns correct data). */
static void render_ray(const vto_volume *vol, const vto_octree *oct, const vto_camera *cam,
                       const vto_params *p, int64_t id, trav_stack *st,
                       float *px, vto_hit *h, uint8_t *brick_seen)
{
    int i = (int)(id / cam->W), j = (int)(id % cam->W);
    float d[3];
    vto_ray_dir(cam, i, j, d);
    memset(h, 0, sizeof *h);
    px[0] = px[1] = px[2] = px[3] = 1.0f;   /* NOCUBE -> white, rayCaster.cu:146-153 */
    trav_ctx t;
    trav_init(&t, oct, cam->pos, d);
    if (!root_ok(&t)) return;
    trav_reset(st);
    int n = oct->nLevels;
    int side = 1 << (n - p->levelO);
    int shift = 3 * (p->levelO - p->levelC);
    uint64_t node;
    while (trav_next(&t, st, p->levelO, &node)) {
        int32_t mn[3];
        vto_decode(node, p->levelO, n, mn);
        if (brick_seen) {
            uint64_t b = node >> shift;
            brick_seen[b - (1ull << (3 * p->levelC))] = 1;
        }
        float Xn[3];
        if (march_node(vol, cam->pos, d, mn, side, p->iso, Xn)) {
            shade(vol, Xn, p, px);
            h->node = node;
            h->brick = node >> shift;
            h->voxel[0] = floor_to_int(Xn[0]);
            h->voxel[1] = floor_to_int(Xn[1]);
            h->voxel[2] = floor_to_int(Xn[2]);
            float dx = Xn[0] - cam->pos[0], dy = Xn[1] - cam->pos[1], dz = Xn[2] - cam->pos[2];
            h->depth = sqrtf(dx * dx + dy * dy + dz * dz);
            return;
        }
    }
}

int vto_render(const vto_volume *vol, const vto_octree *oct,
               const vto_camera *cam, const vto_params *p,
               int64_t ray0, int64_t ray1,
               float *rgba, vto_hit *hits, uint8_t *brick_seen, int nthreads)
{
    if (p->levelC > p->levelO || p->levelO > oct->maxLevel) return -1;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        trav_stack *st = (trav_stack *)malloc(sizeof(trav_stack));
        /* 64-ray chunks keep neighbouring pixels on one thread. */
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t c = ray0; c < ray1; c += 64) {
            int64_t e = c + 64 < ray1 ? c + 64 : ray1;
            for (int64_t id = c; id < e; ++id)
                render_ray(vol, oct, cam, p, id, st, rgba + 4 * id, hits + id, brick_seen);
        }
        free(st);
    }
    return 0;
}

int vto_render_list(const vto_volume *vol, const vto_octree *oct,
                    const vto_camera *cam, const vto_params *p,
                    const int64_t *ids, int64_t count,
                    float *rgba, vto_hit *hits, int nthreads)
{
    if (p->levelC > p->levelO || p->levelO > oct->maxLevel) return -1;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        trav_stack *st = (trav_stack *)malloc(sizeof(trav_stack));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t k = 0; k < count; ++k)
            render_ray(vol, oct, cam, p, ids[k], st, rgba + 4 * k, hits + k, NULL);
        free(st);
    }
    return 0;
}

/* ========================================================= cached frame */

/* visualTur.cu:167-228 with its brick cache, for the CPU baseline: every
 * pass advances each waiting ray's DFS to its next node (Octree.cu:883-901,
 * the rays in parallel), then updates an LRU cache of `capacity` bricks in
 * ray order (lruCache.cu:168-313: a hit moves the brick to the back of the
 * list, a miss takes the front brick's slot and reads the zero-padded box of
 * the missing brick from the source, lruCache.cu:226, at most `capacity`
 * new bricks per pass), then marches every ray whose brick is cached
 * (rayCaster.cu:139-265, in parallel) from the slot.  One policy departs
 * from the reference: a slot whose brick a ray of the current pass uses is
 * never taken (the reference's GPU tier could overwrite it, Q2); a ray
 * whose brick finds no slot waits for the next pass.  Pixels and hit
 * records are those of vto_render. */
#define VTO_CSTACK 48   /* a line meets at most 4 children per level: 3 net pushes per level */

static void lru_unlink(int32_t *prev, int32_t *next, int32_t *head, int32_t *tail, int32_t s)
{
    if (prev[s] >= 0) next[prev[s]] = next[s]; else *head = next[s];
    if (next[s] >= 0) prev[next[s]] = prev[s]; else *tail = prev[s];
    prev[s] = next[s] = -1;
}

static void lru_push_back(int32_t *prev, int32_t *next, int32_t *head, int32_t *tail, int32_t s)
{
    prev[s] = *tail;
    next[s] = -1;
    if (*tail >= 0) next[*tail] = s; else *head = s;
    *tail = s;
}

struct vto_cache {
    int level_c, cube, inc, S;
    int64_t capacity, nb, pass;   /* pass: stamp of the current pass, across frames */
    size_t slot_elems;
    float *pool;
    int32_t *slot_of, *prev, *next;
    int64_t *brick_of, *used_in;
    int32_t head, tail;
};

vto_cache *vto_cache_new(int level_c, int cube, int inc, int64_t capacity)
{
    if (level_c < 0 || level_c > 10 || cube < 1 || inc < 0 || capacity < 1) return NULL;
    vto_cache *c = (vto_cache *)calloc(1, sizeof *c);
    if (!c) return NULL;
    c->level_c = level_c;
    c->cube = cube;
    c->inc = inc;
    c->S = cube + 2 * inc;
    c->nb = (int64_t)1 << (3 * level_c);
    c->capacity = capacity < c->nb ? capacity : c->nb;
    c->slot_elems = (size_t)c->S * c->S * c->S;
    c->pool = (float *)malloc((size_t)c->capacity * c->slot_elems * sizeof(float));
    c->slot_of = (int32_t *)malloc((size_t)c->nb * sizeof(int32_t));
    c->brick_of = (int64_t *)malloc((size_t)c->capacity * sizeof(int64_t));
    c->used_in = (int64_t *)malloc((size_t)c->capacity * sizeof(int64_t));
    c->prev = (int32_t *)malloc((size_t)c->capacity * sizeof(int32_t));
    c->next = (int32_t *)malloc((size_t)c->capacity * sizeof(int32_t));
    if (!c->pool || !c->slot_of || !c->brick_of || !c->used_in || !c->prev || !c->next) {
        vto_cache_free(c);
        return NULL;
    }
    for (int64_t b = 0; b < c->nb; ++b) c->slot_of[b] = -1;
    c->head = c->tail = -1;
    for (int64_t s = 0; s < c->capacity; ++s) {   /* every slot starts empty, in the list */
        c->brick_of[s] = -1;
        c->used_in[s] = -1;
        lru_push_back(c->prev, c->next, &c->head, &c->tail, (int32_t)s);
    }
    return c;
}

void vto_cache_free(vto_cache *c)
{
    if (!c) return;
    free(c->pool); free(c->slot_of); free(c->brick_of); free(c->used_in); free(c->prev); free(c->next);
    free(c);
}

int vto_render_cached(const vto_octree *oct, const vto_camera *cam, const vto_params *p, vto_cache *c,
                      vto_box_reader read, void *user, float *rgba, vto_hit *hits, vto_cache_stats *stats,
                      int nthreads)
{
    if (!c || p->levelC != c->level_c || p->levelC > p->levelO || p->levelO > oct->maxLevel) return -1;
    const int n = oct->nLevels;
    if ((1 << (n - p->levelC)) != c->cube) return -1;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
    const int64_t R = (int64_t)cam->W * cam->H;
    const int S = c->S, inc = c->inc;
    const uint64_t base = (uint64_t)1 << (3 * p->levelC);
    const int shift = 3 * (p->levelO - p->levelC);
    const int side = 1 << (n - p->levelO);
    uint8_t *state = (uint8_t *)malloc((size_t)R);        /* 0 advance, 1 wait for its brick, 2 cached, 3 done */
    uint64_t *node = (uint64_t *)malloc((size_t)R * sizeof(uint64_t));
    int32_t *slot = (int32_t *)malloc((size_t)R * sizeof(int32_t));
    float *dirs = (float *)malloc((size_t)R * 3 * sizeof(float));
    uint64_t *sidx = (uint64_t *)malloc((size_t)R * VTO_CSTACK * sizeof(uint64_t));
    int *slvl = (int *)malloc((size_t)R * VTO_CSTACK * sizeof(int));
    int *stop = (int *)malloc((size_t)R * sizeof(int));
    int rc = 0;
    if (!state || !node || !slot || !dirs || !sidx || !slvl || !stop) {
        rc = -2;
        goto done;
    }
    vto_cache_stats st;
    memset(&st, 0, sizeof st);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int64_t id = 0; id < R; ++id) {   /* visualTur.cu:59-81 and Octree.cu:750-760: every ray at the root */
        float *d = dirs + 3 * id;
        vto_ray_dir(cam, (int)(id / cam->W), (int)(id % cam->W), d);
        float *px = rgba + 4 * id;
        memset(&hits[id], 0, sizeof hits[id]);
        px[0] = px[1] = px[2] = px[3] = 1.0f;
        trav_ctx t;
        trav_init(&t, oct, cam->pos, d);
        state[id] = root_ok(&t) ? 0 : 3;
        sidx[id * VTO_CSTACK] = 1;
        slvl[id * VTO_CSTACK] = 0;
        stop[id] = 0;
    }
    for (;;) {
        int64_t waiting = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256) reduction(+ : waiting)
#endif
        for (int64_t id = 0; id < R; ++id) {   /* getBoxIntersected */
            if (state[id] == 0) {
                trav_ctx t;
                trav_init(&t, oct, cam->pos, dirs + 3 * id);
                uint64_t nd;
                if (trav_next_s(&t, sidx + id * VTO_CSTACK, slvl + id * VTO_CSTACK, &stop[id], VTO_CSTACK,
                                p->levelO, &nd)) {
                    node[id] = nd;
                    state[id] = 1;
                } else {
                    state[id] = 3;   /* exhausted: stays white */
                }
            }
            if (state[id] == 1) waiting++;
        }
        if (waiting == 0) break;
        st.passes++;
        const int64_t pass = ++c->pass;
        /* updateCache, in ray order */
        int64_t loaded = 0;
        for (int64_t id = 0; id < R; ++id) {
            if (state[id] != 1) continue;
            const int64_t b = (int64_t)((node[id] >> shift) - base);
            int32_t sl = c->slot_of[b];
            if (sl >= 0) {
                st.hits++;
            } else {
                if (loaded >= c->capacity) continue;   /* the pass's admission is used up: wait */
                int32_t v = c->head;   /* least recently used slot no ray of this pass uses */
                while (v >= 0 && c->used_in[v] == pass) v = c->next[v];
                if (v < 0) continue;   /* every slot serves this pass: wait */
                if (c->brick_of[v] >= 0) {
                    c->slot_of[c->brick_of[v]] = -1;
                    st.evictions++;
                }
                int32_t mn[3];
                vto_decode((uint64_t)b + base, p->levelC, n, mn);
                int32_t s3[3] = {mn[0] - inc, mn[1] - inc, mn[2] - inc};
                int32_t e3[3] = {s3[0] + S, s3[1] + S, s3[2] + S};
                if (read(user, s3, e3, c->pool + (size_t)v * c->slot_elems) != 0) {
                    c->brick_of[v] = -1;
                    rc = -3;
                    goto done;
                }
                c->brick_of[v] = b;
                c->slot_of[b] = v;
                sl = v;
                loaded++;
                st.misses++;
            }
            c->used_in[sl] = pass;
            lru_unlink(c->prev, c->next, &c->head, &c->tail, sl);
            lru_push_back(c->prev, c->next, &c->head, &c->tail, sl);
            slot[id] = sl;
            state[id] = 2;
        }
        /* render: march the cached rays from their slots */
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
#endif
        for (int64_t id = 0; id < R; ++id) {
            if (state[id] != 2) continue;
            const uint64_t b = node[id] >> shift;
            int32_t bmn[3], nmn[3];
            vto_decode(b, p->levelC, n, bmn);
            vto_decode(node[id], p->levelO, n, nmn);
            vto_volume sv;
            sv.data = c->pool + (size_t)slot[id] * c->slot_elems;
            sv.dims[0] = sv.dims[1] = sv.dims[2] = S;
            sv.origin[0] = bmn[0] - inc;
            sv.origin[1] = bmn[1] - inc;
            sv.origin[2] = bmn[2] - inc;
            const float *d = dirs + 3 * id;
            float Xn[3];
            if (march_node(&sv, cam->pos, d, nmn, side, p->iso, Xn)) {
                vto_hit *h = &hits[id];
                shade(&sv, Xn, p, rgba + 4 * id);
                h->node = node[id];
                h->brick = b;
                h->voxel[0] = floor_to_int(Xn[0]);
                h->voxel[1] = floor_to_int(Xn[1]);
                h->voxel[2] = floor_to_int(Xn[2]);
                float dx = Xn[0] - cam->pos[0], dy = Xn[1] - cam->pos[1], dz = Xn[2] - cam->pos[2];
                h->depth = sqrtf(dx * dx + dy * dy + dz * dz);
                state[id] = 3;
            } else {
                state[id] = 0;
            }
        }
    }
    if (stats) *stats = st;
done:
    free(state); free(node); free(slot); free(dirs); free(sidx); free(slvl); free(stop);
    return rc;
}

/* FileManager.cu:86-167 on an in-memory volume, as a brick source. */
int vto_volume_box_reader(void *user, const int32_t s[3], const int32_t e[3], float *out)
{
    vto_read_box((const vto_volume *)user, s, e, out);
    return 0;
}

/* rayCaster.cu:267-357 (cuda_rayCasterCube) for every pixel: one brick with
 * interior [minBox, minBox+dim)^3, sampled from the zero-padded volume
 * (the brick is a box read of it).  The step counter is a float and the
 * step count is (tfar-tnear)/0.5 without rounding up; colour uses dim. */
int vto_render_cube(const vto_volume *vol, const vto_camera *cam, float iso, const float light[3],
                    const int32_t minBox[3], int dim, float *rgba)
{
    vto_params p;
    memset(&p, 0, sizeof p);
    p.iso = iso;
    p.light[0] = light[0]; p.light[1] = light[1]; p.light[2] = light[2];
    p.color_scale = (float)dim;
    int64_t R = (int64_t)cam->W * cam->H;
    for (int64_t id = 0; id < R; ++id) {
        float d[3], *px = rgba + 4 * id;
        vto_ray_dir(cam, (int)(id / cam->W), (int)(id % cam->W), d);
        const float *o = cam->pos;
        int32_t mx[3] = {minBox[0] + dim, minBox[1] + dim, minBox[2] + dim};
        float tnear, tfar;
        int hit = 0;
        float Xnew[3] = {0, 0, 0};
        if (vto_slab(o, d, minBox, mx, &tnear, &tfar)) {
            float X[3], Xfar[3], vStep[3];
            for (int k = 0; k < 3; ++k) X[k] = o[k] + tnear * d[k];
            for (int k = 0; k < 3; ++k) { Xfar[k] = X[k]; Xnew[k] = X[k]; }
            for (int k = 0; k < 3; ++k) vStep[k] = 0.5f * d[k];
            float maxStep = (float)((double)(tfar - tnear) / 0.5);
            if (maxStep > (float)(4 * dim + 8)) maxStep = (float)(4 * dim + 8);
            int first = 1;
            float ant = 0.0f, sig = 0.0f, steps = 0.0f;
            while (steps <= maxStep) {
                if (first) {
                    first = 0;
                    ant = vto_sample(vol, X[0], X[1], X[2]);
                    for (int k = 0; k < 3; ++k) Xfar[k] = X[k];
                } else {
                    sig = vto_sample(vol, X[0], X[1], X[2]);
                    if (((iso - ant) < 0.0 && (iso - sig) < 0.0) || ((iso - ant) > 0.0 && (iso - sig) > 0.0)) {
                        ant = sig;
                        for (int k = 0; k < 3; ++k) Xfar[k] = X[k];
                    } else {
                        float a = (iso - ant) / (sig - ant);
                        for (int k = 0; k < 3; ++k) Xnew[k] = Xfar[k] * (1.0f - a) + X[k] * a;
                        hit = 1;
                        steps = maxStep;
                    }
                }
                for (int k = 0; k < 3; ++k) X[k] = X[k] + vStep[k];
                steps++;
            }
        }
        if (hit) shade(vol, Xnew, &p, px);
        else px[0] = px[1] = px[2] = px[3] = 1.0f;
    }
    return 0;
}

/* =============================================================== builder */

typedef struct {
    uint64_t *e;
    size_t n, cap;
    uint64_t next;
    int started;
} rle_level;

static void rle_push(rle_level *L, uint64_t v)
{
    if (L->n == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 64;
        L->e = (uint64_t *)realloc(L->e, L->cap * sizeof(uint64_t));
    }
    L->e[L->n++] = v;
}

/* CreateOctree.cu:79-116 (OctreeLevel::insert). */
static void rle_insert(rle_level *L, uint64_t idx)
{
    if (!L->started) {
        rle_push(L, idx);
        L->started = 1;
        L->next = idx + 1;
    } else if (L->next == idx) {
        L->next++;
    } else {
        rle_push(L, L->next - 1);
        rle_push(L, idx);
        L->next = idx + 1;
    }
}

/* CreateOctree.cu:118-128 (completeLevel). */
static void rle_complete(rle_level *L)
{
    if (L->started) rle_push(L, L->next - 1);
}

/* CreateOctree.cu:300-334 (OctreeM ctor) and 425-573 (createTree). */
int vto_build_octree(const vto_volume *vol, float iso, int literal,
                     uint64_t **runs_out, int32_t *sizes_out, int max_levels)
{
    int dimension;
    const int32_t *rd = vol->dims;
    if (rd[0] > rd[1] && rd[0] > rd[2]) dimension = rd[0];
    else if (rd[1] > rd[2]) dimension = rd[1];
    else dimension = rd[2];
    float aux = logf((float)dimension) / logf(2.0f);
    float aux2 = aux - floorf(aux);
    int n = aux2 > 0.0 ? (int)(aux + 1) : (int)aux;
    if (n + 1 > max_levels || n > 20) return -1;
    int D = 1 << n;

    rle_level *lv = (rle_level *)calloc((size_t)n + 1, sizeof(rle_level));
    /* leaf level: every cell in Morton order, 8-corner sign test
     * (CreateOctree.cu:452-511). */
    uint64_t inicio = vto_encode(0, 0, 0, n, n);
    uint64_t fin = vto_encode(D - 1, D - 1, D - 1, n, n);
    for (uint64_t c = inicio; c <= fin; ++c) {
        int32_t m[3];
        vto_decode(c, n, n, m);
        float vx[8];
        vx[0] = vto_voxel(vol, m[0], m[1], m[2]);
        vx[1] = vto_voxel(vol, m[0], m[1], m[2] + 1);
        vx[2] = vto_voxel(vol, m[0], m[1] + 1, m[2]);
        vx[3] = vto_voxel(vol, m[0], m[1] + 1, m[2] + 1);
        vx[4] = vto_voxel(vol, m[0] + 1, m[1], m[2]);
        vx[5] = vto_voxel(vol, m[0] + 1, m[1], m[2] + 1);
        vx[6] = vto_voxel(vol, m[0] + 1, m[1] + 1, m[2]);
        vx[7] = vto_voxel(vol, m[0] + 1, m[1] + 1, m[2] + 1);
        int has = 0;
        int sign = (vx[0] - iso) < 0;
        for (int i = 1; i < 8 && !has; ++i)
            if (((vx[i] - iso) < 0) != sign) has = 1;
        if (has) rle_insert(&lv[n], c);
    }
    rle_complete(&lv[n]);
    /* inner levels (CreateOctree.cu:541-571); literal mode keeps the
     * `inicio != fin` bound that never visits the last node. */
    for (int level = n - 1; level >= 1; --level) {
        uint64_t a = vto_encode(0, 0, 0, level, n);
        uint64_t b = vto_encode(D - 1, D - 1, D - 1, level, n);
        uint64_t end = literal ? b : b + 1;
        int csize = (int)lv[level + 1].n;
        for (uint64_t c = a; c != end; ++c)
            if (vto_children_mask_builder(lv[level + 1].e, csize, c) != 0)
                rle_insert(&lv[level], c);
        rle_complete(&lv[level]);
    }
    /* root (CreateOctree.cu:532-540) */
    if (n >= 1 && vto_children_mask_builder(lv[1].e, (int)lv[1].n, 1) != 0) {
        rle_insert(&lv[0], 1);
        rle_complete(&lv[0]);
    }
    if (n == 0) {   /* single-cell volume: level 0 is the leaf level */
        /* handled by the leaf pass above (inicio == fin == 1) */
    }
    size_t total = 0;
    for (int l = 0; l <= n; ++l) total += lv[l].n;
    uint64_t *runs = (uint64_t *)malloc((total ? total : 1) * sizeof(uint64_t));
    size_t off = 0;
    for (int l = 0; l <= n; ++l) {
        sizes_out[l] = (int32_t)lv[l].n;
        if (lv[l].n) memcpy(runs + off, lv[l].e, lv[l].n * sizeof(uint64_t));
        off += lv[l].n;
        free(lv[l].e);
    }
    free(lv);
    *runs_out = runs;
    return n;
}

void vto_free(void *p) { free(p); }

/* CreateOctree.cu:342-371 (writeToFile). */
int vto_octree_write(const char *path, float iso, int32_t dimension,
                     const int32_t realDim[3], int32_t nLevels,
                     const uint64_t *runs, const int32_t *sizes)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    int32_t magic = VTO_MAGIC;
    fwrite(&magic, 4, 1, f);
    fwrite(&iso, 4, 1, f);
    fwrite(&dimension, 4, 1, f);
    fwrite(realDim, 4, 3, f);
    fwrite(&nLevels, 4, 1, f);
    size_t *offs = (size_t *)malloc(((size_t)nLevels + 1) * sizeof(size_t));
    size_t off = 0;
    for (int l = 0; l <= nLevels; ++l) { offs[l] = off; off += (size_t)sizes[l]; }
    for (int l = nLevels; l >= 0; --l) {
        int32_t ne = sizes[l];
        fwrite(&ne, 4, 1, f);
        if (ne) fwrite(runs + offs[l], 8, (size_t)ne, f);
    }
    free(offs);
    return fclose(f) == 0 ? 0 : -1;
}

/* Octree.cu:775-815 (the file-reading part of the Octree ctor). */
int vto_octree_read(const char *path, float *iso, int32_t *dimension,
                    int32_t realDim[3], int32_t *nLevels,
                    uint64_t **runs_out, int32_t *sizes_out, int max_levels)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int32_t magic = 0;
    if (fread(&magic, 4, 1, f) != 1 || magic != VTO_MAGIC) { fclose(f); return -2; }
    if (fread(iso, 4, 1, f) != 1 || fread(dimension, 4, 1, f) != 1 ||
        fread(realDim, 4, 3, f) != 3 || fread(nLevels, 4, 1, f) != 1) { fclose(f); return -3; }
    int n = *nLevels;
    if (n < 0 || n + 1 > max_levels) { fclose(f); return -4; }
    if (!runs_out) { fclose(f); return 0; }
    uint64_t **lv = (uint64_t **)calloc((size_t)n + 1, sizeof(uint64_t *));
    size_t total = 0;
    int rc = 0;
    for (int l = n; l >= 0; --l) {
        int32_t ne = 0;
        if (fread(&ne, 4, 1, f) != 1 || ne < 0) { rc = -5; break; }
        sizes_out[l] = ne;
        lv[l] = (uint64_t *)malloc(((size_t)ne + 1) * sizeof(uint64_t));
        if (ne && fread(lv[l], 8, (size_t)ne, f) != (size_t)ne) { rc = -5; break; }
        total += (size_t)ne;
    }
    fclose(f);
    if (rc == 0) {
        uint64_t *runs = (uint64_t *)malloc((total ? total : 1) * sizeof(uint64_t));
        size_t off = 0;
        for (int l = 0; l <= n; ++l) {
            if (sizes_out[l]) memcpy(runs + off, lv[l], (size_t)sizes_out[l] * 8);
            off += (size_t)sizes_out[l];
        }
        *runs_out = runs;
    }
    for (int l = 0; l <= n; ++l) free(lv[l]);
    free(lv);
    return rc;
}
