// vt_device.hpp -- device-side building blocks of the isosurface hot path.
//
// Everything here is plain FP32 with no contraction (the file sets
// `fp contract(off)` and the build passes -ffp-contract=off); division and
// sqrtf are the correctly rounded HIP defaults.  The arithmetic order of each
// routine follows the reference line for line so that node codes, brick ids
// and hit voxels come out bit-identical to the reference algorithm.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace vt {

// ------------------------------------------------------------------ Morton
// A node code at level l is a 1 sentinel bit followed by l (x,y,z) bit
// triples, x highest (mortonCodeUtil.hpp:125-148).

__host__ __device__ __forceinline__ uint32_t compact3(uint64_t x)
{
    // keep every third bit (bit 3b -> bit b), inverse of the bit dilation
    x &= 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x001f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x001f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x00000000001fffffull;
    return (uint32_t)x;
}

// compact3 for codes below 2^30 (nodes of levels <= 9 without their
// leading 1 bit), in 32-bit arithmetic: half the instructions
__host__ __device__ __forceinline__ uint32_t compact3_30(uint32_t x)
{
    x &= 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030c30c3u;
    x = (x ^ (x >> 4)) & 0x0300f00fu;
    x = (x ^ (x >> 8)) & 0x030000ffu;
    x = (x ^ (x >> 16)) & 0x000003ffu;
    return x;
}

__host__ __device__ __forceinline__ uint64_t dilate3(uint32_t v)
{
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x001f00000000ffffull;
    x = (x | (x << 16)) & 0x001f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// level of a node code (code >= 1), mortonCodeUtil.hpp:21-65
__device__ __forceinline__ int code_level(uint64_t c)
{
    return (63 - __clzll((long long)c)) / 3;
}

// min corner (voxels) of node `c` at level `l` in a 2^n cube,
// mortonCodeUtil.hpp:99-123
__host__ __device__ __forceinline__ void node_min(uint64_t c, int l, int n, int& x, int& y, int& z)
{
    uint64_t m = c ^ (1ull << (3 * l));
    int s = n - l;
    x = (int)(compact3(m >> 2) << s);
    y = (int)(compact3(m >> 1) << s);
    z = (int)(compact3(m) << s);
}

// node_min for levels l <= 9 (the octree's levelO and everything above it)
__host__ __device__ __forceinline__ void node_min9(uint64_t c, int l, int n, int& x, int& y, int& z)
{
    const uint32_t m = (uint32_t)c ^ (1u << (3 * l));
    const int s = n - l;
    x = (int)(compact3_30(m >> 2) << s);
    y = (int)(compact3_30(m >> 1) << s);
    z = (int)(compact3_30(m) << s);
}

// ------------------------------------------------------------------- rays

struct Camera {
    float pos[3], look[3], up[3], right[3];
    float h, w, distance;
    int W, H;
};

struct Ray {
    float ox, oy, oz;
    float dx, dy, dz;
    float ix, iy, iz;   // 1/d, computed once (Octree.cu:268 recomputes the same value)
};

// Camera.cu:58-73; normalize as v * (1/sqrt(dot(v,v)))
__device__ __forceinline__ void make_ray(const Camera& c, int i, int j, Ray& r)
{
    float ih = c.h / (float)c.H;
    float iw = c.w / (float)c.W;
    float su = -(c.h / 2.0f) + (ih * ((float)i + 0.5f));
    float sr = -(c.w / 2.0f) + (iw * ((float)j + 0.5f));
    float ax = c.look[0] * c.distance, ay = c.look[1] * c.distance, az = c.look[2] * c.distance;
    ax = ax + c.up[0] * su; ay = ay + c.up[1] * su; az = az + c.up[2] * su;
    ax = ax + c.right[0] * sr; ay = ay + c.right[1] * sr; az = az + c.right[2] * sr;
    float inv = 1.0f / sqrtf(ax * ax + ay * ay + az * az);
    r.dx = ax * inv; r.dy = ay * inv; r.dz = az * inv;
    r.ox = c.pos[0]; r.oy = c.pos[1]; r.oz = c.pos[2];
    r.ix = 1 / r.dx; r.iy = 1 / r.dy; r.iz = 1 / r.dz;
}

// ------------------------------------------------------------------- slab
// Octree.cu:258-330 (_cuda_RayAABB2) == rayCaster.cu:12-79, box [mn, mx]
__device__ __forceinline__ bool slab_box(const Ray& r, int mnx, int mny, int mnz, int mxx, int mxy, int mxz,
                                         float& tnear, float& tfar)
{
    bool hit = true;
    float tmin, tmax, tymin, tymax, tzmin, tzmax;
    float fx0 = (float)mnx, fx1 = (float)mxx;
    float fy0 = (float)mny, fy1 = (float)mxy;
    float fz0 = (float)mnz, fz1 = (float)mxz;
    if (r.ix >= 0) { tmin = (fx0 - r.ox) * r.ix; tmax = (fx1 - r.ox) * r.ix; }
    else           { tmin = (fx1 - r.ox) * r.ix; tmax = (fx0 - r.ox) * r.ix; }
    if (r.iy >= 0) { tymin = (fy0 - r.oy) * r.iy; tymax = (fy1 - r.oy) * r.iy; }
    else           { tymin = (fy1 - r.oy) * r.iy; tymax = (fy0 - r.oy) * r.iy; }
    if ((tmin > tymax) || (tymin > tmax)) hit = false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    if (r.iz >= 0) { tzmin = (fz0 - r.oz) * r.iz; tzmax = (fz1 - r.oz) * r.iz; }
    else           { tzmin = (fz1 - r.oz) * r.iz; tzmax = (fz0 - r.oz) * r.iz; }
    if ((tmin > tzmax) || (tzmin > tmax)) hit = false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    tnear = (tmin < 0.0f) ? 0.0f : tmin;
    tfar = tmax;
    return hit;
}

__device__ __forceinline__ bool slab(const Ray& r, int mnx, int mny, int mnz, int side, float& tnear, float& tfar)
{
    return slab_box(r, mnx, mny, mnz, mnx + side, mny + side, mnz + side, tnear, tfar);
}

// --------------------------------------------------------------- octree
// Membership is held as one byte per parent: byte (parent - 8^(l-1)) of the
// level-l table has bit k set when child (parent<<3)+k is in the level-l run
// list.  Built on the host from the reference run lists.

struct OctreeView {
    const uint8_t* child_bytes;   // levels 1..maxLevel concatenated (level l: 8^(l-1) bytes)
    uint64_t level_off[24];       // byte offset of level l's table, (8^(l-1) - 1) / 7
    int n;                        // cube side 2^n
    int final_level;              // levelO: the level the traversal emits
    int root_valid;               // level-0 run list contains node 1
};

// The tables follow each other from level 1, so the offset of level
// plevel + 1 is (8^plevel - 1) / 7, computed here rather than looked up in
// level_off by a lane-varying index (which would be one more dependent load
// per walk step).  Parent levels are below 9, so 32 bits hold every index.
__device__ __forceinline__ uint32_t child_mask(const OctreeView& o, uint64_t parent, int plevel)
{
    const uint32_t p8 = 1u << (3 * plevel);
    return o.child_bytes[(uint32_t)parent - p8 + (p8 - 1u) / 7u];
}

// Literal restatement of the bubble sort of Octree.cu:665-706 on <= 8 keys.
// Only used when a key is NaN (degenerate rays); otherwise the rank form
// below gives the identical stable order.
// Keys arrive by value so that the common path keeps them in registers.
__device__ __noinline__ uint32_t bubble_order(float n0, float n1, float n2, float n3, float n4, float n5,
                                              float n6, float n7, float f0, float f1, float f2, float f3,
                                              float f4, float f5, float f6, float f7, uint32_t valid)
{
    const float tn_in[8] = {n0, n1, n2, n3, n4, n5, n6, n7};
    const float tf_in[8] = {f0, f1, f2, f3, f4, f5, f6, f7};
    float tn[8], tf[8];
    int ks[8];
    int nv = 0;
    for (int k = 0; k < 8; ++k)
        if ((valid >> k) & 1) { tn[nv] = tn_in[k]; tf[nv] = tf_in[k]; ks[nv] = k; nv++; }
    int n = nv;
    while (n != 0) {
        int newn = 0;
        for (int id = 1; id < nv; ++id) {
            bool sw = (tn[id - 1] > tn[id]) || (tn[id - 1] == tn[id] && tf[id - 1] > tf[id]);
            if (sw) {
                int a = ks[id]; ks[id] = ks[id - 1]; ks[id - 1] = a;
                float f = tn[id]; tn[id] = tn[id - 1]; tn[id - 1] = f;
                f = tf[id]; tf[id] = tf[id - 1]; tf[id - 1] = f;
                newn = id;
            }
        }
        n = newn;
    }
    uint32_t order = (uint32_t)nv << 24;
    for (int p = 0; p < nv; ++p) order |= (uint32_t)ks[p] << (3 * p);
    return order;
}

// Children of `parent` that are members, slab-hit and have tfar >= 0,
// in the order the reference pushes/pops them (Octree.cu:375-532, 663-712).
// Packed result: bits [3p,3p+3) = child k at position p, bits [24,28) = count.
__device__ __forceinline__ uint32_t sorted_children(const Ray& r, const OctreeView& o,
                                                    uint64_t parent, int plevel)
{
    uint32_t members = child_mask(o, parent, plevel);
    if (members == 0) return 0;
    int px, py, pz;
    node_min9(parent, plevel, o.n, px, py, pz);
    int side = 1 << (o.n - plevel - 1);
    // The children's faces lie on three planes per axis (parent min, middle,
    // max).  slab_box computes (plane - origin) * inverse per face, so these
    // nine products are exactly its per-face values; each child's slab is
    // then slab_box's comparisons on the two planes that bound it per axis.
    float tx[3], ty[3], tz[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        tx[q] = ((float)(px + q * side) - r.ox) * r.ix;
        ty[q] = ((float)(py + q * side) - r.oy) * r.iy;
        tz[q] = ((float)(pz + q * side) - r.oz) * r.iz;
    }
    float tn[8], tf[8];
    uint32_t valid = 0;
    bool has_nan = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        tn[k] = 0.0f; tf[k] = 0.0f;
        if ((members >> k) & 1) {
            const int bx = (k >> 2) & 1, by = (k >> 1) & 1, bz = k & 1;
            float tmin = r.ix >= 0 ? tx[bx] : tx[bx + 1], tmax = r.ix >= 0 ? tx[bx + 1] : tx[bx];
            float tymin = r.iy >= 0 ? ty[by] : ty[by + 1], tymax = r.iy >= 0 ? ty[by + 1] : ty[by];
            float tzmin = r.iz >= 0 ? tz[bz] : tz[bz + 1], tzmax = r.iz >= 0 ? tz[bz + 1] : tz[bz];
            bool hit = true;
            if ((tmin > tymax) || (tymin > tmax)) hit = false;
            if (tymin > tmin) tmin = tymin;
            if (tymax < tmax) tmax = tymax;
            if ((tmin > tzmax) || (tzmin > tmax)) hit = false;
            if (tzmin > tmin) tmin = tzmin;
            if (tzmax < tmax) tmax = tzmax;
            float a = (tmin < 0.0f) ? 0.0f : tmin, b = tmax;
            if (hit && b >= 0.0f) {
                valid |= 1u << k;
                tn[k] = a; tf[k] = b;
                has_nan |= (a != a);   // b is not NaN here (b >= 0 held)
            }
        }
    }
    if (valid == 0) return 0;
    if (has_nan)
        return bubble_order(tn[0], tn[1], tn[2], tn[3], tn[4], tn[5], tn[6], tn[7], tf[0], tf[1], tf[2], tf[3],
                            tf[4], tf[5], tf[6], tf[7], valid);
    // Without NaN the comparison (tnear, then tfar) is a strict weak order and
    // bubble sort is stable, so position = number of valid children that sort
    // strictly before, plus equal keys with a smaller k.
    uint32_t order = (uint32_t)__popc(valid) << 24;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (!((valid >> k) & 1)) continue;
        int rank = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (q == k || !((valid >> q) & 1)) continue;
            bool before = (tn[q] < tn[k]) || (tn[q] == tn[k] && tf[q] < tf[k]) ||
                          (tn[q] == tn[k] && tf[q] == tf[k] && q < k);
            rank += before ? 1 : 0;
        }
        order |= (uint32_t)k << (3 * rank);
    }
    return order;
}

__device__ __forceinline__ int order_count(uint32_t order) { return (int)(order >> 24); }
__device__ __forceinline__ int order_at(uint32_t order, int p) { return (int)((order >> (3 * p)) & 7u); }

// The sorted child list without sorting.  Along one axis a parent's
// children lie between three planes, whose times are the nine products
// sorted_children computes (near, mid, far in ray order).  The children the
// ray passes through, in time order, are the one holding the ray at
// t0 = max(entry, 0) and then the neighbour across each mid-plane the ray
// crosses before it leaves the parent.  When no two of the times involved
// are equal (and none is NaN), every such child has an interval of positive
// length, no other child has one, their tnear values are distinct, and
// tnear order is time order: exactly the list sorted_children returns.  An
// equality (a ray through a node edge or corner, or an origin on a plane)
// reports !ok and the caller takes the sorted list; so does a step from a
// child with no positive-length interval (only reached through that list).
// Returns the first member child after `after` in that order (after < 0:
// the first of all), or -1.
__device__ __forceinline__ int seq_child(const Ray& r, const OctreeView& o, uint64_t parent, int plevel, int after,
                                         bool& ok)
{
    ok = true;
    const uint32_t members = child_mask(o, parent, plevel);
    if (members == 0) return -1;
    int px, py, pz;
    node_min9(parent, plevel, o.n, px, py, pz);
    const int side = 1 << (o.n - plevel - 1);
    const float x0 = ((float)px - r.ox) * r.ix, x1 = ((float)(px + side) - r.ox) * r.ix,
                x2 = ((float)(px + 2 * side) - r.ox) * r.ix;
    const float y0 = ((float)py - r.oy) * r.iy, y1 = ((float)(py + side) - r.oy) * r.iy,
                y2 = ((float)(py + 2 * side) - r.oy) * r.iy;
    const float z0 = ((float)pz - r.oz) * r.iz, z1 = ((float)(pz + side) - r.oz) * r.iz,
                z2 = ((float)(pz + 2 * side) - r.oz) * r.iz;
    // a product is NaN only as 0 * inf: a zero direction component (its
    // inverse infinite) with the origin on one of the planes
    const bool inf_dir = !(fabsf(r.ix) < __builtin_inff()) || !(fabsf(r.iy) < __builtin_inff()) ||
                         !(fabsf(r.iz) < __builtin_inff());
    if (inf_dir && (x0 != x0 || x1 != x1 || x2 != x2 || y0 != y0 || y1 != y1 || y2 != y2 || z0 != z0 ||
                    z1 != z1 || z2 != z2)) {
        ok = false;
        return -1;
    }
    // ray order per axis, as sorted_children picks tmin/tmax (r.i* >= 0)
    const bool nx_ = !(r.ix >= 0), ny_ = !(r.iy >= 0), nz_ = !(r.iz >= 0);
    const float fx = nx_ ? x0 : x2, fy = ny_ ? y0 : y2, fz = nz_ ? z0 : z2;   // far planes
    int c;
    if (after < 0) {
        const float ex = nx_ ? x2 : x0, ey = ny_ ? y2 : y0, ez = nz_ ? z2 : z0;   // near planes
        float t0 = fmaxf(fmaxf(ex, ey), ez);
        t0 = t0 < 0.0f ? 0.0f : t0;
        const float t1 = fminf(fminf(fx, fy), fz);
        if (x1 == t0 || y1 == t0 || z1 == t0 || t1 == t0) {
            ok = false;
            return -1;
        }
        if (t1 < t0) return -1;   // the parent lies behind the origin
        // a crossed mid-plane puts the ray in the far half (in ray order)
        c = ((int)((x1 < t0) != nx_) << 2) | ((int)((y1 < t0) != ny_) << 1) | (int)((z1 < t0) != nz_);
        if ((members >> c) & 1) return c;
    } else {
        // the child the walk stands on must lie on the ray for a positive
        // time (after t = 0): one reached through the sorted list at an
        // equality may only touch it, and the order after it is the list's
        c = after;
        const float ex = nx_ ? x2 : x0, ey = ny_ ? y2 : y0, ez = nz_ ? z2 : z0;
        const bool hx = ((c >> 2) & 1) == (int)nx_, hy = ((c >> 1) & 1) == (int)ny_, hz = (c & 1) == (int)nz_;
        const float tn = fmaxf(fmaxf(fmaxf(hx ? ex : x1, hy ? ey : y1), hz ? ez : z1), 0.0f);
        const float tf = fminf(fminf(hx ? x1 : fx, hy ? y1 : fy), hz ? z1 : fz);
        if (!(tn < tf)) {
            ok = false;
            return -1;
        }
    }
    for (int q = 0; q < 3; ++q) {
        // the child's exit: its far plane per axis (the mid-plane when the
        // child is the near half in ray order)
        const bool hx = ((c >> 2) & 1) == (int)nx_, hy = ((c >> 1) & 1) == (int)ny_, hz = (c & 1) == (int)nz_;
        const float ox = hx ? x1 : fx, oy = hy ? y1 : fy, oz = hz ? z1 : fz;
        const float m = fminf(fminf(ox, oy), oz);
        if ((int)(ox == m) + (int)(oy == m) + (int)(oz == m) > 1) {
            ok = false;
            return -1;
        }
        if (ox == m) {
            if (!hx) return -1;   // leaves the parent
            c ^= 4;
        } else if (oy == m) {
            if (!hy) return -1;
            c ^= 2;
        } else {
            if (!hz) return -1;
            c ^= 1;
        }
        if ((members >> c) & 1) return c;
    }
    return -1;
}

// Next node of the reference's DFS emission sequence at level final_level.
//   descend = true : `node` itself has not been explored yet
//   descend = false: `node` (and its subtree) is finished
// Returns 0 when the traversal is exhausted.  This is the stack DFS of
// Octree.cu:620-720 without the stack: the parent of a node is code>>3 and
// the next child after it is found from the parent's planes (seq_child, in
// ray order), or from the parent's sorted child list where seq_child finds
// equal times; both are pure functions of (parent, ray).
__device__ __forceinline__ uint64_t next_node_seq(const Ray& r, const OctreeView& o, uint64_t node, bool descend)
{
    int level = code_level(node);
    for (;;) {
        if (descend && level == o.final_level) return node;
        if (!descend && level == 0) return 0;
        const uint64_t parent = descend ? node : node >> 3;
        const int plevel = descend ? level : level - 1;
        const int k = descend ? -1 : (int)(node & 7u);
        bool ok;
        int c = seq_child(r, o, parent, plevel, k, ok);
        if (!ok) {
            const uint32_t ord = sorted_children(r, o, parent, plevel);
            const int cnt = order_count(ord);
            int p = 0;
            if (k >= 0) {
                while (p < cnt && order_at(ord, p) != k) ++p;
                p += 1;
            }
            c = p < cnt ? order_at(ord, p) : -1;
        }
        if (c >= 0) {
            node = (parent << 3) | (uint64_t)c;
            level = plevel + 1;
            descend = true;
        } else if (descend) {
            descend = false;   // no child to walk: the node is done
        } else {
            node = parent;
            level = plevel;
        }
    }
}

// Root acceptance of Octree.cu:729: level 0 holds node 1 and the ray meets
// the cube (no tfar test at the root).
__device__ __forceinline__ bool root_ok(const Ray& r, const OctreeView& o)
{
    if (!o.root_valid) return false;
    float a, b;
    return slab(r, 0, 0, 0, 1 << o.n, a, b);
}

// ------------------------------------------------------------ sampling

__device__ __forceinline__ bool coord_ok(float p)
{
    return fabsf(p) < 16777216.0f;   // finite and exactly integral-convertible
}

// rayCaster.cu:86-106 (getElementInterpolate) on a brick slot of side S
// whose voxel (0,0,0) is global voxel (bx,by,bz).  A read outside the slot
// sets `err` and yields 0; a non-finite coordinate yields NaN.
__device__ __forceinline__ float sample(const float* __restrict__ slot, int S, int bx, int by, int bz,
                                        float px, float py, float pz, uint32_t& err)
{
    if (!coord_ok(px) || !coord_ok(py) || !coord_ok(pz)) return __builtin_nanf("");
    float ix, iy, iz;
    float fx = modff(px, &ix);
    float fy = modff(py, &iy);
    float fz = modff(pz, &iz);
    int x0 = (int)ix - bx, y0 = (int)iy - by, z0 = (int)iz - bz;
    if ((unsigned)x0 >= (unsigned)(S - 1) || (unsigned)y0 >= (unsigned)(S - 1) ||
        (unsigned)z0 >= (unsigned)(S - 1)) {
        err |= 1u;
        return 0.0f;
    }
    const float* p = slot + ((size_t)x0 * S + y0) * S + z0;
    size_t sx = (size_t)S * S, sy = (size_t)S;
    float v000 = p[0], v001 = p[1], v010 = p[sy], v011 = p[sy + 1];
    float v100 = p[sx], v101 = p[sx + 1], v110 = p[sx + sy], v111 = p[sx + sy + 1];
    float c00 = v000 * (1.0f - fx) + v100 * fx;
    float c01 = v001 * (1.0f - fx) + v101 * fx;
    float c10 = v010 * (1.0f - fx) + v110 * fx;
    float c11 = v011 * (1.0f - fx) + v111 * fx;
    float c0 = c00 * (1.0f - fy) + c10 * fy;
    float c1 = c01 * (1.0f - fy) + c11 * fy;
    return c0 * (1.0f - fz) + c1 * fz;
}

__device__ __forceinline__ int cvt_rzi(float v)
{
    // truncate with saturation, NaN -> 0 (the hardware conversion semantics)
    if (v != v) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return (int)0x80000000;
    return (int)v;
}

// rayCaster.cu:156-237 for one node whose brick slot is `slot`.
// Returns true on a crossing with the secant point in X*.
__device__ __forceinline__ bool march_node(const Ray& r, int nx, int ny, int nz, int side,
                                           const float* __restrict__ slot, int S,
                                           int bx, int by, int bz, float iso,
                                           float& Xx, float& Xy, float& Xz,
                                           uint32_t& err, uint32_t& nsamples)
{
    float tnear, tfar;
    if (!slab(r, nx, ny, nz, side, tnear, tfar)) return false;
    float X0 = r.ox + tnear * r.dx, X1 = r.oy + tnear * r.dy, X2 = r.oz + tnear * r.dz;
    float F0 = X0, F1 = X1, F2 = X2;
    float s0 = 0.5f * r.dx, s1 = 0.5f * r.dy, s2 = 0.5f * r.dz;
    int maxStep = cvt_rzi(ceilf((tfar - tnear) * 2.0f));
    if (maxStep > 4 * side + 8) maxStep = 4 * side + 8;
    if (maxStep < 0) return false;
    float ant = sample(slot, S, bx, by, bz, X0, X1, X2, err);
    nsamples += 1;
    for (int steps = 1; steps <= maxStep; ++steps) {
        X0 = X0 + s0; X1 = X1 + s1; X2 = X2 + s2;
        float sig = sample(slot, S, bx, by, bz, X0, X1, X2, err);
        nsamples += 1;
        if (((iso - ant) < 0.0f && (iso - sig) < 0.0f) || ((iso - ant) > 0.0f && (iso - sig) > 0.0f)) {
            ant = sig;
            F0 = X0; F1 = X1; F2 = X2;
        } else {
            float a = (iso - ant) / (sig - ant);
            Xx = F0 * (1.0f - a) + X0 * a;
            Xy = F1 * (1.0f - a) + X1 * a;
            Xz = F2 * (1.0f - a) + X2 * a;
            return true;
        }
    }
    return false;
}

// ----------------------------------------------- march with sample skipping
// Same sample sequence, positions, crossing test and secant as march_node,
// but a sample is not evaluated when the min/max of the 4x4x4-cell
// sub-block holding its cell proves that the interpolated value lies
// strictly on one side of iso: then only its side matters to the crossing
// test.  The margin 2^-16 * max(|lo|,|hi|,|iso|) is far above the rounding
// error of the trilinear form (convex weights for non-negative positions),
// so the proven side is the side the float evaluation would give.  A value
// the secant needs (the last sample before a crossing) is evaluated exactly.

enum Side : int { SIDE_NONE = 0, SIDE_ABOVE = 1, SIDE_BELOW = 2 };   // value above / below iso

__device__ __forceinline__ int side_of(float iso, float v)
{
    return (iso - v) < 0.0f ? SIDE_ABOVE : ((iso - v) > 0.0f ? SIDE_BELOW : SIDE_NONE);
}

struct SubGrid {
    const float2* mm;   // G^3 (min, max) of this brick, x-major, or null
    int G;              // sub-blocks per axis (cube / 4)
    int cx0, cy0, cz0;  // global voxel of the brick's first core cell
    // the sides side_from_minmax gives for this pass's iso, 2 bits per
    // sub-block (sub-block sb in bits 2*(sb & 15) of word sb >> 4), or null:
    // one 128-byte line per 32^3 brick instead of a 4 KB (min, max) grid
    const uint32_t* sides = nullptr;
};

// Side of iso that a sub-block's (min, max) proves for every trilinear value
// inside it, with the margin 2^-16 * max(|lo|, |hi|, |iso|).
__device__ __forceinline__ int side_from_minmax(float lo, float hi, float iso)
{
    float M = fmaxf(fmaxf(fabsf(lo), fabsf(hi)), fabsf(iso));
    float m = M * 1.52587890625e-05f + 1e-30f;   // 2^-16 * M
    if (lo > iso && (lo - iso) > m) return 1;     // SIDE_ABOVE
    if (hi < iso && (iso - hi) > m) return 2;     // SIDE_BELOW
    return 0;
}

// The sub-block last looked up: its voxel box [x0,x0+4) x [y0,y0+4) x
// [z0,z0+4) (integer-valued floats, so x0 + 4 is exact) and its proven side.
// For p >= 0, (int)p is floor(p), so p lies in the box exactly when its
// cell lies in the sub-block; NaN fails every comparison (the empty box).
struct SideCache {
    float x0 = __builtin_nanf(""), y0 = __builtin_nanf(""), z0 = __builtin_nanf("");
    int side = SIDE_NONE;
};

__device__ __forceinline__ int proven_side(const SubGrid& g, float iso, float px, float py, float pz, SideCache& sc)
{
    if (px >= sc.x0 && px < sc.x0 + 4.0f && py >= sc.y0 && py < sc.y0 + 4.0f && pz >= sc.z0 && pz < sc.z0 + 4.0f)
        return sc.side;
    if ((!g.mm && !g.sides) || !(px >= 0.0f) || !(py >= 0.0f) || !(pz >= 0.0f) || !coord_ok(px) || !coord_ok(py) ||
        !coord_ok(pz))
        return SIDE_NONE;
    int cx = (int)px - g.cx0, cy = (int)py - g.cy0, cz = (int)pz - g.cz0;
    int lim = 4 * g.G;
    if ((unsigned)cx >= (unsigned)lim || (unsigned)cy >= (unsigned)lim || (unsigned)cz >= (unsigned)lim)
        return SIDE_NONE;
    int sb = ((cx >> 2) * g.G + (cy >> 2)) * g.G + (cz >> 2);
    int s;
    if (g.sides) {
        s = (int)((g.sides[sb >> 4] >> (2 * (sb & 15))) & 3u);
    } else {
        float2 mm = g.mm[sb];
        s = side_from_minmax(mm.x, mm.y, iso);
    }
    sc.x0 = (float)(g.cx0 + (cx & ~3));
    sc.y0 = (float)(g.cy0 + (cy & ~3));
    sc.z0 = (float)(g.cz0 + (cz & ~3));
    sc.side = s;
    return s;
}

// Steps, at most `lim`, that certainly keep the position inside the cached
// sub-block box when taken from X by the accumulation X += s (each add
// rounded).  After m adds X_a differs from X_a + m s_a by at most m half-ulps
// of values below |X_a| + 8; a ray crosses a 4-voxel box in at most 14 steps
// of half a voxel, so a margin of 16 * 2^-22 * (|X| + 8) + 1e-5 voxel covers
// the drift and the rounding of this estimate.  s = d / 2, so 1 / |s_a| is
// 2 |1 / d_a|, from the ray's inverse directions.
__device__ __forceinline__ int steps_inside(const SideCache& sc, const Ray& r, float X0, float X1, float X2, float s0,
                                            float s1, float s2, int lim)
{
    const float inv0 = 2.0f * fabsf(r.ix), inv1 = 2.0f * fabsf(r.iy), inv2 = 2.0f * fabsf(r.iz);
    const float m = fmaxf(fmaxf(fabsf(X0), fabsf(X1)), fabsf(X2)) + 8.0f;
    const float margin = m * 3.814697265625e-06f + 1e-5f;
    float kf = (float)lim;
    if (s0 != 0.0f) kf = fminf(kf, ((s0 > 0.0f ? (sc.x0 + 4.0f) - X0 : X0 - sc.x0) - margin) * inv0);
    if (s1 != 0.0f) kf = fminf(kf, ((s1 > 0.0f ? (sc.y0 + 4.0f) - X1 : X1 - sc.y0) - margin) * inv1);
    if (s2 != 0.0f) kf = fminf(kf, ((s2 > 0.0f ? (sc.z0 + 4.0f) - X2 : X2 - sc.z0) - margin) * inv2);
    return kf >= 1.0f ? (int)kf : 0;   // NaN gives 0
}

__device__ __forceinline__ bool march_node_skip(const Ray& r, int nx, int ny, int nz, int side,
                                                const float* __restrict__ slot, int S, int bx, int by, int bz,
                                                const SubGrid& g, float iso, float& Xx, float& Xy, float& Xz,
                                                uint32_t& err, uint32_t& nsamples)
{
    float tnear, tfar;
    if (!slab(r, nx, ny, nz, side, tnear, tfar)) return false;
    float X0 = r.ox + tnear * r.dx, X1 = r.oy + tnear * r.dy, X2 = r.oz + tnear * r.dz;
    const float s0 = 0.5f * r.dx, s1 = 0.5f * r.dy, s2 = 0.5f * r.dz;
    int maxStep = cvt_rzi(ceilf((tfar - tnear) * 2.0f));
    if (maxStep > 4 * side + 8) maxStep = 4 * side + 8;
    if (maxStep < 0) return false;
    SideCache sc;
    // the previous sample: its position F*, its value ant (valid only when
    // ant_known) and its side
    float F0 = X0, F1 = X1, F2 = X2;
    float ant = 0.0f;
    bool ant_known;
    int ant_side = proven_side(g, iso, X0, X1, X2, sc);
    if (ant_side == SIDE_NONE) {
        ant = sample(slot, S, bx, by, bz, X0, X1, X2, err);
        nsamples += 1;
        ant_known = true;
        ant_side = side_of(iso, ant);
    } else {
        ant_known = false;
    }
    for (int steps = 1; steps <= maxStep; ++steps) {
        X0 = X0 + s0; X1 = X1 + s1; X2 = X2 + s2;
        int sd = proven_side(g, iso, X0, X1, X2, sc);
        if (sd != SIDE_NONE && sd == ant_side) {   // same strict side: the reference's "ant = sig"
            // the next steps inside the same sub-block take this branch too:
            // accumulate their positions without the box test
            const int k = steps_inside(sc, r, X0, X1, X2, s0, s1, s2, maxStep - steps);
            for (int q = 0; q < k; ++q) {
                X0 = X0 + s0; X1 = X1 + s1; X2 = X2 + s2;
            }
            steps += k;
            F0 = X0; F1 = X1; F2 = X2;
            ant_known = false;
            continue;
        }
        float sig = sample(slot, S, bx, by, bz, X0, X1, X2, err);
        nsamples += 1;
        int sig_side = side_of(iso, sig);
        if (sig_side != SIDE_NONE && sig_side == ant_side) {
            ant = sig;
            ant_known = true;
            F0 = X0; F1 = X1; F2 = X2;
            continue;
        }
        if (!ant_known) {   // the secant needs the previous value itself
            ant = sample(slot, S, bx, by, bz, F0, F1, F2, err);
            nsamples += 1;
        }
        float a = (iso - ant) / (sig - ant);
        Xx = F0 * (1.0f - a) + X0 * a;
        Xy = F1 * (1.0f - a) + X1 * a;
        Xz = F2 * (1.0f - a) + X2 * a;
        return true;
    }
    return false;
}

// rayCaster.cu:238-251 from the central-difference gradient (n0, n1, n2)
__device__ __forceinline__ void shade_normal(float n0, float n1, float n2, float Xx, float Xy, float Xz, float lx,
                                             float ly, float lz, float color_scale, float4& out)
{
    float ninv = 1.0f / sqrtf(n0 * n0 + n1 * n1 + n2 * n2);
    n0 = n0 * ninv; n1 = n1 * ninv; n2 = n2 * ninv;
    float l0 = Xx - lx, l1 = Xy - ly, l2 = Xz - lz;
    float linv = 1.0f / sqrtf(l0 * l0 + l1 * l1 + l2 * l2);
    l0 = l0 * linv; l1 = l1 * linv; l2 = l2 * linv;
    float dif = fabsf(n0 * l0 + n1 * l1 + n2 * l2);
    float a = Xy / color_scale;
    out.x = (1 - a) * dif;
    out.y = (a) * dif;
    out.z = 0.0f * dif;
    out.w = 1.0f;
}

// rayCaster.cu:131-137 and 238-251
__device__ __forceinline__ void shade(const float* __restrict__ slot, int S, int bx, int by, int bz,
                                      float Xx, float Xy, float Xz,
                                      float lx, float ly, float lz, float color_scale,
                                      float4& out, uint32_t& err)
{
    // the six central-difference samples in one loop (one inlined sample
    // site): step q takes axis q>>1, at -1 for even q and +1 for odd q;
    // n_axis = (s(-1) - s(+1)) / 2 exactly as written out per axis
    float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f, lo = 0.0f;
#pragma unroll 1
    for (int q = 0; q < 6; ++q) {
        const int ax = q >> 1;
        const bool plus = (q & 1) != 0;
        const float px = ax == 0 ? (plus ? Xx + 1.0f : Xx - 1.0f) : Xx;
        const float py = ax == 1 ? (plus ? Xy + 1.0f : Xy - 1.0f) : Xy;
        const float pz = ax == 2 ? (plus ? Xz + 1.0f : Xz - 1.0f) : Xz;
        const float s = sample(slot, S, bx, by, bz, px, py, pz, err);
        if (!plus) {
            lo = s;
        } else {
            const float v = (lo - s) / 2.0f;
            n0 = ax == 0 ? v : n0;
            n1 = ax == 1 ? v : n1;
            n2 = ax == 2 ? v : n2;
        }
    }
    shade_normal(n0, n1, n2, Xx, Xy, Xz, lx, ly, lz, color_scale, out);
}

}  // namespace vt
