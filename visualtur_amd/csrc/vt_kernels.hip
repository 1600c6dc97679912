// vt_kernels.hip -- gfx950 kernels of the isosurface hot path.
//
// k_render_pass fuses what the reference runs as three stages per pass
// (Octree::getBoxIntersected, lruCache::updateCache, rayCaster::render,
// visualTur.cu:175-220): one thread per ray walks the octree in the
// reference's front-to-back order, marches every emitted node whose brick is
// resident, shades the first crossing and stops.  A ray whose next brick is
// not resident records the brick in the request set and parks; the host
// loads the requested bricks and relaunches over the parked rays only.
#include <algorithm>

#include "vt_internal.hpp"

#pragma clang fp contract(off)

namespace vt {

namespace {

constexpr int kBlock = 256;   // 4 wave64s = one 16x16 pixel tile

// Wave-aggregated atomic add on a 32-bit counter: one atomic per wave.
__device__ __forceinline__ uint32_t wave_atomic_add(uint32_t* ctr, bool pred)
{
    unsigned long long mask = __ballot(pred);
    if (!pred) return 0;
    int lane = (int)__lane_id();
    unsigned long long lower = mask & ((1ull << lane) - 1ull);
    int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(ctr, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(lower);
}

__device__ __forceinline__ int floor_to_int(float x) { return cvt_rzi(floorf(x)); }

}  // namespace

__device__ __forceinline__ void set_miss(float4& out, HitRecord& h)
{
    out = make_float4(1.0f, 1.0f, 1.0f, 1.0f);   // rayCaster.cu:146-153
    h.node = 0; h.brick = 0; h.voxel[0] = h.voxel[1] = h.voxel[2] = 0; h.depth = 0.0f;
}

// Marches node `node` of brick `brick`, resident in `slot`; on a crossing,
// shades it into out/h and returns true.
__device__ __forceinline__ bool visit_node(const Ray& r, const OctreeView& oct, const BrickView& bv, const Shading& sh,
                                           uint64_t node, uint64_t brick, int slot, uint8_t* slot_ref,
                                           uint8_t* brick_trace, float4& out, HitRecord& h, uint32_t& err,
                                           uint32_t& nsamples, uint32_t& nnodes)
{
    const int n = oct.n, levelO = oct.final_level;
    const int side = 1 << (n - levelO);
    if (slot_ref) slot_ref[slot] = 1;
    if (brick_trace) brick_trace[brick - bv.brick_base] = 1;
    int nx, ny, nz, bx, by, bz;
    node_min9(node, levelO, n, nx, ny, nz);
    node_min9(brick, bv.level_c, n, bx, by, bz);
    bx -= bv.inc; by -= bv.inc; bz -= bv.inc;
    const float* sp = bv.pool + (size_t)slot * bv.slot_elems;
    float Xx, Xy, Xz;
    nnodes += 1;
    // without sub-block min/max (mm == null) no side is ever proven and
    // march_node_skip evaluates every sample: march_node's sequence exactly
    SubGrid g{bv.submm ? bv.submm + (size_t)slot * bv.G * bv.G * bv.G : nullptr, bv.G, bx + bv.inc,
              by + bv.inc, bz + bv.inc, bv.sides ? bv.sides + (size_t)slot * bv.side_words : nullptr};
    if (!march_node_skip(r, nx, ny, nz, side, sp, bv.S, bx, by, bz, g, sh.iso, Xx, Xy, Xz, err, nsamples))
        return false;
    shade(sp, bv.S, bx, by, bz, Xx, Xy, Xz, sh.lx, sh.ly, sh.lz, sh.color_scale, out, err);
    nsamples += 6;
    h.node = node;
    h.brick = brick;
    h.voxel[0] = floor_to_int(Xx);
    h.voxel[1] = floor_to_int(Xy);
    h.voxel[2] = floor_to_int(Xz);
    float dx = Xx - r.ox, dy = Xy - r.oy, dz = Xz - r.oz;
    h.depth = sqrtf(dx * dx + dy * dy + dz * dz);
    return true;
}

// Adds brick `bidx` to this pass's request set.  The first request appends
// bidx + 1 to the list in coherent host memory with a system-scope store,
// so the host can load the brick while this pass is still running.
__device__ __forceinline__ void request_brick(const RenderArgs& a, uint64_t bidx)
{
    uint32_t bit = 1u << (bidx & 31);
    uint32_t old = atomicOr(&a.req_bits[bidx >> 5], bit);
    if (!(old & bit)) {
        uint32_t q = atomicAdd(&a.ctr->req_count, 1u);
        __hip_atomic_store(&a.req_list[q], (uint32_t)bidx + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ------------------------------------------------------ wide march
// In the passes after the first only the parked rays run, a few thousand or
// fewer, and each walks alone through a chain of dependent loads while the
// GPU idles.  The wide pass gives each ray a group of L lanes (L = 2^lg,
// groups aligned inside the wave) that march every node together: lane q of
// the group takes samples q, q+L, q+2L, ... of the node, so the sub-block
// lookups and trilinear gathers of L samples are in flight at once, and a
// group ballot finds the first crossing.  Everything else of the ray (the
// traversal, the brick lookups, the deferral list) is computed identically
// by every lane of the group; only the group's first lane writes or issues
// atomics.
struct Lanes {
    int L;         // lanes per ray (1 = the one-lane march)
    int sub;       // this lane's index in its group
    int base;      // wave lane of the group's first lane
    bool leader;   // sub == 0
};

__device__ __forceinline__ uint64_t group_ballot(bool p, const Lanes& g)
{
    const unsigned long long m = __ballot(p);
    return g.L == 64 ? (uint64_t)m : ((uint64_t)m >> g.base) & ((1ull << g.L) - 1ull);
}

// march_node_skip's sample sequence, crossing test and secant, marched by
// the L lanes of a group.  Positions are accumulated serially by every lane
// (X_k = X_(k-1) + step, as the reference does); lane q keeps X_k for its k.
// A sample is evaluated when its sub-block proves no side; the side of
// sample k is then the proven side or the side of the evaluated value, and
// the crossing is the first k >= 1 whose side is not the strict side of
// sample k-1.  `err` gets exactly the flags of the samples the one-lane march
// evaluates (every unproven sample up to the crossing, the crossing sample,
// and the sample before it when that was skipped); `nsamples` gets their
// count (group-uniform).  Results are group-uniform.
__device__ __forceinline__ bool march_node_wide(const Ray& r, int nx, int ny, int nz, int side,
                                                const float* __restrict__ slot, int S, int bx, int by, int bz,
                                                const SubGrid& sg, float iso, float& Xx, float& Xy, float& Xz,
                                                uint32_t& err, uint32_t& nsamples, const Lanes& g)
{
    float tnear, tfar;
    if (!slab(r, nx, ny, nz, side, tnear, tfar)) return false;
    float R0 = r.ox + tnear * r.dx, R1 = r.oy + tnear * r.dy, R2 = r.oz + tnear * r.dz;
    const float s0 = 0.5f * r.dx, s1 = 0.5f * r.dy, s2 = 0.5f * r.dz;
    int maxStep = cvt_rzi(ceilf((tfar - tnear) * 2.0f));
    if (maxStep > 4 * side + 8) maxStep = 4 * side + 8;
    if (maxStep < 0) return false;
    SideCache sc;
    // the previous chunk's last sample: side, value (if evaluated), position
    int c_side = SIDE_NONE;
    float c_val = 0.0f;
    bool c_known = false;
    float C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    uint32_t cnt = 0;
    const int last = g.base + g.L - 1;
    for (int k0 = 0; k0 <= maxStep; k0 += g.L) {
        float P0 = R0, P1 = R1, P2 = R2;
        for (int j = 0; j < g.L; ++j) {
            if (j == g.sub) { P0 = R0; P1 = R1; P2 = R2; }
            R0 = R0 + s0; R1 = R1 + s1; R2 = R2 + s2;
        }
        const int k = k0 + g.sub;
        const bool valid = k <= maxStep;
        const int ps = valid ? proven_side(sg, iso, P0, P1, P2, sc) : (int)SIDE_NONE;
        const bool ev = valid && ps == SIDE_NONE;
        uint32_t e = 0;
        const float v = ev ? sample(slot, S, bx, by, bz, P0, P1, P2, e) : 0.0f;
        const int sd = ps != SIDE_NONE ? ps : (ev ? side_of(iso, v) : (int)SIDE_NONE);
        const int up = __shfl(sd, (g.base + g.sub + 63) & 63);
        const int prev = g.sub == 0 ? c_side : up;
        const bool cross = valid && k >= 1 && !(sd != SIDE_NONE && sd == prev);
        const uint64_t cm = group_ballot(cross, g);
        const uint64_t evm = group_ballot(ev, g);
        if (cm == 0) {
            cnt += (uint32_t)__popcll(evm);
            err |= e;
            c_side = __shfl(sd, last);
            c_val = __shfl(v, last);
            c_known = __shfl((int)ev, last) != 0;
            C0 = __shfl(P0, last);
            C1 = __shfl(P1, last);
            C2 = __shfl(P2, last);
            continue;
        }
        const int js = __ffsll((unsigned long long)cm) - 1;   // the crossing is sample k0 + js
        if (g.sub <= js) err |= e;
        cnt += (uint32_t)__popcll(evm & ((1ull << js) - 1ull)) + 1u;   // unproven before it, and the crossing sample
        // the crossing sample's value
        float sv = v;
        if (g.sub == js && !ev) {
            uint32_t e2 = 0;
            sv = sample(slot, S, bx, by, bz, P0, P1, P2, e2);
            err |= e2;
        }
        const float sig = __shfl(sv, g.base + js);
        const float X0 = __shfl(P0, g.base + js), X1 = __shfl(P1, g.base + js), X2 = __shfl(P2, g.base + js);
        // the previous sample's value and position
        float ant, F0, F1, F2;
        if (js > 0) {
            const bool known = (evm >> (js - 1)) & 1ull;
            float av = v;
            if (g.sub == js - 1 && !known) {
                uint32_t e3 = 0;
                av = sample(slot, S, bx, by, bz, P0, P1, P2, e3);
                err |= e3;
            }
            ant = __shfl(av, g.base + js - 1);
            F0 = __shfl(P0, g.base + js - 1);
            F1 = __shfl(P1, g.base + js - 1);
            F2 = __shfl(P2, g.base + js - 1);
            cnt += known ? 0u : 1u;
        } else {
            F0 = C0; F1 = C1; F2 = C2;
            if (c_known) {
                ant = c_val;
            } else {
                uint32_t e4 = 0;
                ant = sample(slot, S, bx, by, bz, F0, F1, F2, e4);
                err |= e4;
                cnt += 1u;
            }
        }
        const float a = (iso - ant) / (sig - ant);
        Xx = F0 * (1.0f - a) + X0 * a;
        Xy = F1 * (1.0f - a) + X1 * a;
        Xz = F2 * (1.0f - a) + X2 * a;
        nsamples += cnt;
        return true;
    }
    nsamples += cnt;
    return false;
}

// shade() with the six central-difference samples taken by six lanes of the
// group (L >= 8); the arithmetic after the samples is shade()'s.
__device__ __forceinline__ void shade_wide(const float* __restrict__ slot, int S, int bx, int by, int bz, float Xx,
                                           float Xy, float Xz, float lx, float ly, float lz, float color_scale,
                                           float4& out, uint32_t& err, const Lanes& g)
{
    const int q = g.sub < 6 ? g.sub : 0;
    const int ax = q >> 1;
    const bool plus = (q & 1) != 0;
    const float px = ax == 0 ? (plus ? Xx + 1.0f : Xx - 1.0f) : Xx;
    const float py = ax == 1 ? (plus ? Xy + 1.0f : Xy - 1.0f) : Xy;
    const float pz = ax == 2 ? (plus ? Xz + 1.0f : Xz - 1.0f) : Xz;
    uint32_t e = 0;
    const float sv = g.sub < 6 ? sample(slot, S, bx, by, bz, px, py, pz, e) : 0.0f;
    err |= e;
    const float v0 = __shfl(sv, g.base + 0), v1 = __shfl(sv, g.base + 1);
    const float v2 = __shfl(sv, g.base + 2), v3 = __shfl(sv, g.base + 3);
    const float v4 = __shfl(sv, g.base + 4), v5 = __shfl(sv, g.base + 5);
    shade_normal((v0 - v1) / 2.0f, (v2 - v3) / 2.0f, (v4 - v5) / 2.0f, Xx, Xy, Xz, lx, ly, lz, color_scale, out);
}

// visit_node for a group of lanes (kWide) or one lane.
template <bool kWide>
__device__ __forceinline__ bool visit_node_g(const Ray& r, const OctreeView& oct, const BrickView& bv,
                                             const Shading& sh, uint64_t node, uint64_t brick, int slot,
                                             uint8_t* slot_ref, uint8_t* brick_trace, float4& out, HitRecord& h,
                                             uint32_t& err, uint32_t& nsamples, uint32_t& nnodes, const Lanes& g)
{
    if (!kWide)
        return visit_node(r, oct, bv, sh, node, brick, slot, slot_ref, brick_trace, out, h, err, nsamples, nnodes);
    const int n = oct.n, levelO = oct.final_level;
    const int side = 1 << (n - levelO);
    if (g.leader) {
        if (slot_ref) slot_ref[slot] = 1;
        if (brick_trace) brick_trace[brick - bv.brick_base] = 1;
    }
    int nx, ny, nz, bx, by, bz;
    node_min9(node, levelO, n, nx, ny, nz);
    node_min9(brick, bv.level_c, n, bx, by, bz);
    bx -= bv.inc; by -= bv.inc; bz -= bv.inc;
    const float* sp = bv.pool + (size_t)slot * bv.slot_elems;
    float Xx, Xy, Xz;
    nnodes += 1;
    SubGrid sg{bv.submm ? bv.submm + (size_t)slot * bv.G * bv.G * bv.G : nullptr, bv.G, bx + bv.inc,
               by + bv.inc, bz + bv.inc, bv.sides ? bv.sides + (size_t)slot * bv.side_words : nullptr};
    if (!march_node_wide(r, nx, ny, nz, side, sp, bv.S, bx, by, bz, sg, sh.iso, Xx, Xy, Xz, err, nsamples, g))
        return false;
    if (g.L >= 8)
        shade_wide(sp, bv.S, bx, by, bz, Xx, Xy, Xz, sh.lx, sh.ly, sh.lz, sh.color_scale, out, err, g);
    else
        shade(sp, bv.S, bx, by, bz, Xx, Xy, Xz, sh.lx, sh.ly, sh.lz, sh.color_scale, out, err);
    nsamples += 6;
    h.node = node;
    h.brick = brick;
    h.voxel[0] = floor_to_int(Xx);
    h.voxel[1] = floor_to_int(Xy);
    h.voxel[2] = floor_to_int(Xz);
    float dx = Xx - r.ox, dy = Xy - r.oy, dz = Xz - r.oz;
    h.depth = sqrtf(dx * dx + dy * dy + dz * dz);
    return true;
}

// Render-pass workgroup: kPassBlock/64 waves, each rendering one 8x8 pixel
// tile (the tiles of a workgroup are consecutive in row-major tile order).
constexpr int kPassBlock = 64;

// One ray of a render pass (ray `id`, when `active`; group-uniform).  With
// kMulti the ray belongs to frame slot `fslot` (wave-uniform; id = fslot <<
// a.rshift | pixel; its camera, shading and outputs in a.jobs), `first` says
// whether this is its frame's first pass, and the per-ray state is indexed
// by id with deferral lists of a.defer_stride rays; otherwise the frame is
// a.cam and the first pass is the one without an input list.

template <bool kWide, bool kMulti>
__device__ __forceinline__ void pass_ray(const RenderArgs& a, uint32_t id, bool active, bool first_region,
                                         const Lanes& g, uint32_t fslot)
{
    const uint32_t pix = kMulti ? (id & ((1u << a.rshift) - 1u)) : id;
    const Camera& cam = kMulti ? a.jobs[fslot].cam : a.cam;
    const int W = cam.W, H = cam.H;
    const uint32_t R = kMulti ? a.defer_stride : (uint32_t)W * (uint32_t)H;
    const bool first = kMulti ? first_region : a.in_list == nullptr;
    const int max_defer = kMulti && !first ? a.max_defer_late : a.max_defer;
    const int node_budget = kMulti && !first ? 0 : a.node_budget;
    uint32_t err = 0, nsamples = 0, nnodes = 0;
    bool park = false;

    if (active) {
        int i = (int)(pix / (uint32_t)W), j = (int)(pix % (uint32_t)W);
        Ray r;
        make_ray(cam, i, j, r);
        const Shading sh = kMulti ? a.jobs[fslot].sh : Shading{a.iso, a.lx, a.ly, a.lz, a.color_scale};
        // the proven-side words hold the sides for a.side_iso (never NaN)
        BrickView bv{a.brick_slot, a.pool, a.S, a.inc, a.slot_elems, a.level_c, a.shift, a.brick_base, a.submm, a.G,
                     a.sides && sh.iso == a.side_iso ? a.sides : nullptr, a.side_words};
        float4 c;
        HitRecord h;
        uint64_t bidx = 0, node = 0;
        int nd = 0;
        // what the pass leaves: WALK_HIT/WALK_MISS = c/h to write, WALK_ABSENT =
        // park on `node`, kKeep = the output already holds the ray's result
        // (final when nd == 0), kBlocked = a deferred node is still absent
        constexpr int kKeep = 3, kBlocked = 4, kBudget = 5;
        int walked = 0;   // walk nodes taken in this pass
        int st;
        // One loop over the nodes this pass marches, so that the march, the
        // shading and the traversal step are each inlined once: first the
        // deferred nodes of earlier passes (k < nprev, in walk order), then
        // the walk itself (from the root in the first pass, from the parked
        // node in later ones; `adv` = advance past `node` before using it).
        // A node whose brick is absent is deferred (its brick requested)
        // while the list has room, and the walk goes on past it: the first
        // crossing in a deferred node, marched in a later pass, precedes any
        // crossing found beyond it, so a walk that ends with nd > 0 leaves a
        // tentative result.
        int nprev = 0, k = 0;
        uint8_t tent = 0;
        bool adv, descend;
        if (first) {
            adv = descend = root_ok(r, a.oct);
            node = adv ? 1ull : 0ull;
        } else {
            const uint8_t ph = a.ray_phase[id];
            nprev = ph & PHASE_NMASK;
            tent = ph & PHASE_TENT;
            adv = descend = false;
            // parked on this node, not yet marched (read only when the walk
            // resumes: a ray with a tentative result has no resume node)
            if (!tent) node = a.ray_node[id];
        }
        for (;;) {
            const bool deferred = k < nprev;
            uint64_t cur;
            if (deferred) {
                cur = a.ray_defer[(size_t)k * R + id];
            } else {
                if (tent) {   // every deferred node missed: the tentative result stands
                    st = kKeep;
                    break;
                }
                if (adv) {
                    node = next_node_seq(r, a.oct, node, descend);
                    adv = descend = false;
                }
                if (node == 0) {
                    set_miss(c, h);
                    st = WALK_MISS;
                    break;
                }
                // a ray that has walked its budget parks on its next node,
                // requesting nothing, and goes on in a later pass
                if (node_budget > 0 && walked >= node_budget) {
                    st = kBudget;
                    break;
                }
                walked += 1;
                cur = node;
            }
            const uint64_t brick = cur >> bv.shift;
            bidx = brick - bv.brick_base;
            const int slot = bv.brick_slot[bidx];
            if (slot < 0) {
                if (deferred) {   // still absent: keep it and the ones after it
                    if (g.leader) {
                        for (int q = k; q < nprev; ++q)
                            a.ray_defer[(size_t)(q - k) * R + id] = a.ray_defer[(size_t)q * R + id];
                        request_brick(a, bidx);
                    }
                    nd = nprev - k;
                    if (g.leader) a.ray_phase[id] = (uint8_t)(tent | nd);   // resume node and tentative result kept
                    park = true;
                    st = kBlocked;
                    break;
                }
                if (nd >= max_defer) {
                    st = WALK_ABSENT;
                    break;
                }
                if (g.leader) {
                    a.ray_defer[(size_t)nd * R + id] = cur;
                    request_brick(a, bidx);
                }
                nd += 1;
            } else if (visit_node_g<kWide>(r, a.oct, bv, sh, cur, brick, slot, a.slot_ref, a.brick_trace, c, h,
                                           err, nsamples, nnodes, g)) {
                st = WALK_HIT;   // final when found in a deferred node (nd == 0 there)
                break;
            }
            if (deferred) {
                k += 1;
            } else {
                adv = true;
            }
        }
        if (g.leader) {
            if (st == WALK_ABSENT || st == kBudget) {
                if (st == WALK_ABSENT) request_brick(a, bidx);
                a.ray_node[id] = node;   // not yet marched
                a.ray_phase[id] = (uint8_t)nd;
                park = true;
            } else if (st == WALK_HIT || st == WALK_MISS) {
                (kMulti ? a.jobs[fslot].rgba : a.rgba)[pix] = c;
                (kMulti ? a.jobs[fslot].hits : a.hits)[pix] = h;
                if (nd > 0) {
                    a.ray_phase[id] = (uint8_t)(PHASE_TENT | nd);
                    park = true;
                }
            }
        }
    }

    // parked rays: one atomic per wave (a multi-frame wave holds rays of one
    // frame, which park in that frame's list)
    park = park && g.leader;
    if (kMulti) {
        const uint32_t pslot = wave_atomic_add(&a.job_parked[fslot], park);
        if (park) a.out_list[(fslot << a.rshift) + pslot] = id;
    } else {
        const uint32_t pslot = wave_atomic_add(&a.ctr->parked_count, park);
        if (park) a.out_list[pslot] = id;
    }
    if (err) atomicOr(&a.ctr->err, err);
    if (a.collect_stats) {
        unsigned long long s = g.leader ? nsamples : 0u, nn = g.leader ? nnodes : 0u;
        for (int off = 32; off > 0; off >>= 1) {
            s += __shfl_xor(s, off);
            nn += __shfl_xor(nn, off);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&a.ctr->samples, s);
            atomicAdd(&a.ctr->nodes, nn);
        }
    }
}


// at most 96 VGPRs: 5 waves per SIMD instead of 4, without spilling
__global__ __launch_bounds__(kPassBlock, 5) void k_render_pass(RenderArgs a)
{
    const int W = a.cam.W, H = a.cam.H;
    uint32_t id;
    bool active = true;
    if (a.in_list) {
        uint32_t idx = blockIdx.x * kPassBlock + threadIdx.x;
        active = idx < a.in_count;
        id = active ? a.in_list[idx] : 0;
    } else {
        // Workgroups take tiles in order and are dealt round-robin to the 8
        // XCDs, so neighbouring tiles land on different XCDs and the
        // expensive rays, which cluster in the image, spread over all CUs.
        // (Giving each XCD a contiguous band of tiles, for L2 reuse, measured
        // 8% slower on the 1024^3 workload: the bands holding the surface
        // became the tail.)  With supertile S > 1 the XCDs are dealt S x S
        // blocks of tiles instead (block q to XCD q % 8, the k-th workgroup
        // of an XCD taking tile k % S^2 of its (k / S^2)-th block), so the
        // tiles of a block share one L2 while the blocks of every XCD still
        // spread over the image.
        const int tx8 = (W + 7) >> 3, ty8 = (H + 7) >> 3;
        const int lane = threadIdx.x & 63;
        int t;
        if (a.supertile > 1) {
            const int S = a.supertile, sx = (tx8 + S - 1) / S, sy = (ty8 + S - 1) / S;
            const int k = (int)(blockIdx.x >> 3), r = k % (S * S);
            const int q = (k / (S * S)) * 8 + (int)(blockIdx.x & 7);
            const int tx = (q % sx) * S + r % S, ty = (q / sx) * S + r / S;
            active = q < sx * sy && tx < tx8 && ty < ty8;
            t = active ? ty * tx8 + tx : 0;
        } else {
            t = (int)blockIdx.x * (kPassBlock / 64) + (int)(threadIdx.x >> 6);
        }
        active = active && t < tx8 * ty8;
        const int i = (active ? t / tx8 : 0) * 8 + (lane >> 3);
        const int j = (active ? t % tx8 : 0) * 8 + (lane & 7);
        active = active && i < H && j < W;
        id = active ? (uint32_t)(i * W + j) : 0;
    }
    const Lanes g{1, 0, (int)(threadIdx.x & 63), true};
    pass_ray<false, false>(a, id, active, false, g, 0u);
}

// A pass with 2^lg lanes per ray over groups of 64 rays: the first pass's
// 8x8 tiles, or 64 consecutive entries of in_list.  A group takes 2^lg
// one-wave workgroups, and they all go to one XCD: workgroup b is wave
// (b/8) mod 2^lg of group ((b/8) >> lg)*8 + b mod 8 (workgroups are dealt
// round-robin to the 8 XCDs), so the group's bricks are fetched into one
// L2, and neighbouring groups still spread over all XCDs.
__global__ __launch_bounds__(kPassBlock) void k_render_wide(RenderArgs a, int lg)
{
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t b = blockIdx.x;
    const uint32_t grp = ((b >> 3) >> lg) * 8u + (b & 7u);
    const uint32_t w = (b >> 3) & ((1u << lg) - 1u);
    const uint32_t p = (w << (6 - lg)) + ((uint32_t)lane >> lg);   // ray of the group, 0..63
    const uint32_t idx = grp * 64u + p;
    bool active;
    uint32_t id;
    if (a.in_list) {
        active = idx < a.in_count;
        id = active ? a.in_list[idx] : 0;
    } else {
        const int W = a.cam.W, H = a.cam.H;
        const int tx8 = (W + 7) >> 3, ty8 = (H + 7) >> 3;
        const int tile = (int)grp;
        active = tile < tx8 * ty8;
        const int i = (active ? tile / tx8 : 0) * 8 + (int)(p >> 3);
        const int j = (active ? tile % tx8 : 0) * 8 + (int)(p & 7);
        active = active && i < H && j < W;
        id = active ? (uint32_t)(i * W + j) : 0;
    }
    const int L = 1 << lg;
    const Lanes g{L, lane & (L - 1), lane & ~(L - 1), (lane & (L - 1)) == 0};
    pass_ray<true, false>(a, id, active, false, g, 0u);
}

// The frame of list group `grp` (64 consecutive entries of one frame's
// parked list; list_block0 holds the prefix sums of the frames' groups) and
// the entry's offset in that list.
__device__ __forceinline__ uint32_t list_frame(const RenderArgs& a, uint32_t grp)
{
    uint32_t s = 0;
    while (s + 1 < (uint32_t)kMaxFrames && grp >= a.list_block0[s + 1]) ++s;
    return s;
}

// Multi-frame pass, one lane per ray: workgroups [0, list groups) take 64
// consecutive entries of one frame's parked list (the later passes of the
// frames in flight, started first since they are the long walks), the rest
// take the 8x8 tiles of the fresh frames in order (fresh_tile0 prefix sums),
// dealt round-robin to the XCDs as in k_render_pass.  With supertile S > 1
// the list region is padded to a multiple of 8 workgroups and each fresh
// frame's region holds ceil(blocks / 8) * 8 * S^2 workgroups, so that
// workgroup b still runs on XCD b % 8 and every XCD is dealt whole S x S
// blocks of a frame's tiles, as in k_render_pass.  Every wave holds rays of
// one frame, so the frame's camera, shading and outputs are wave-uniform.
__global__ __launch_bounds__(kPassBlock, 5) void k_render_multi(RenderArgs a)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t list_blocks = a.list_block0[kMaxFrames];
    const uint32_t list_region = a.supertile > 1 ? (list_blocks + 7u) & ~7u : list_blocks;
    uint32_t id = 0, fs;
    bool active, first;
    if (blockIdx.x < list_blocks) {
        fs = list_frame(a, blockIdx.x);
        const uint32_t off = (blockIdx.x - a.list_block0[fs]) * 64 + lane;
        active = off < a.list_count[fs];
        id = active ? a.in_list[(fs << a.rshift) + off] : 0u;
        first = false;
    } else if (blockIdx.x < list_region) {   // padding
        fs = 0;
        active = false;
        first = false;
    } else {
        const uint32_t t = blockIdx.x - list_region;
        int k = 0;
        while (k + 1 < a.nfresh && t >= a.fresh_tile0[k + 1]) ++k;
        fs = a.fresh_slot[k];
        const Camera& cam = a.jobs[fs].cam;
        const int tx8 = (cam.W + 7) >> 3, ty8 = (cam.H + 7) >> 3;
        const uint32_t tt = t - a.fresh_tile0[k];
        active = t < a.fresh_tile0[a.nfresh];
        int ti, tj;
        if (a.supertile > 1) {
            const int S = a.supertile, sx = (tx8 + S - 1) / S, sy = (ty8 + S - 1) / S;
            const int kk = (int)(tt >> 3), r = kk % (S * S);
            const int q = (kk / (S * S)) * 8 + (int)(tt & 7u);
            tj = (q % sx) * S + r % S;
            ti = (q / sx) * S + r / S;
            active = active && q < sx * sy && tj < tx8 && ti < ty8;
        } else {
            ti = (int)(tt / (uint32_t)tx8);
            tj = (int)(tt % (uint32_t)tx8);
        }
        const int i = ti * 8 + (int)(lane >> 3);
        const int j = tj * 8 + (int)(lane & 7);
        active = active && i < cam.H && j < cam.W;
        id = active ? (fs << a.rshift) | (uint32_t)(i * cam.W + j) : 0u;
        first = true;
    }
    const Lanes g{1, 0, (int)lane, true};
    pass_ray<false, true>(a, id, active, first, g, fs);
}

// Multi-frame pass over in_list only, 2^lg lanes per ray (the later passes
// of frames in flight when no frame starts), grouped as in k_render_wide.
__global__ __launch_bounds__(kPassBlock, 4) void k_render_multi_wide(RenderArgs a, int lg)
{
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t b = blockIdx.x;
    const uint32_t grp = ((b >> 3) >> lg) * 8u + (b & 7u);
    const uint32_t w = (b >> 3) & ((1u << lg) - 1u);
    const uint32_t p = (w << (6 - lg)) + ((uint32_t)lane >> lg);
    const uint32_t fs = list_frame(a, grp);
    const uint32_t off = (grp - a.list_block0[fs]) * 64u + p;
    const bool active = off < a.list_count[fs];
    const uint32_t id = active ? a.in_list[(fs << a.rshift) + off] : 0u;
    const int L = 1 << lg;
    const Lanes g{L, lane & (L - 1), lane & ~(L - 1), (lane & (L - 1)) == 0};
    pass_ray<true, true>(a, id, active, false, g, fs);
}

// ------------------------------------------------- persistent render pass
// The same per-ray work as k_render_pass, but a fixed grid of workgroups
// loops over the pass's rays: a lane whose ray finishes (or parks) takes a
// new ray when a quarter of its wave is idle, so waves stay full and every
// memory latency serves ~64 rays instead of the few still running in a
// fixed 8x8 tile.  Each iteration marches at most one node per lane.
//
// Work items of a pass: the first pass enumerates 16x16 tiles (256 items
// each, in k_render_pass's pixel order, items outside the image skipped);
// later passes enumerate in_list.  The items are cut into 8 contiguous parts,
// one per XCD (workgroups are dealt to XCDs round-robin, so blockIdx % 8 is
// the home part), each with its own counter; a wave whose part is used up
// takes items from the next part.  Every item is taken exactly once, and a
// wave leaves when all parts are used up and its lanes are idle.
namespace {

enum LaneMode : int { LANE_IDLE = 0, LANE_WALK = 1, LANE_DEFER = 2 };

__device__ __forceinline__ void part_range(uint32_t total, uint32_t unit, int p, uint32_t& beg, uint32_t& end)
{
    uint32_t per = ((total + unit - 1) / unit + 7) / 8 * unit;
    beg = per * (uint32_t)p;
    end = beg + per;
    if (beg > total) beg = total;
    if (end > total) end = total;
}

}  // namespace

__global__ __launch_bounds__(kBlock) void k_render_persistent(RenderArgs a)
{
    const int W = a.cam.W, H = a.cam.H;
    const uint32_t R = (uint32_t)W * (uint32_t)H;
    const bool first = a.in_list == nullptr;
    const uint32_t total = first ? (uint32_t)a.tiles_total * 256u : a.in_count;
    const uint32_t unit = first ? 256u : 1u;
    const int lane = (int)__lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    BrickView bv{a.brick_slot, a.pool, a.S, a.inc, a.slot_elems, a.level_c, a.shift, a.brick_base, a.submm, a.G,
                 a.sides && a.iso == a.side_iso ? a.sides : nullptr, a.side_words};
    Shading sh{a.iso, a.lx, a.ly, a.lz, a.color_scale};

    int part = (int)(blockIdx.x & 7u);   // wave-uniform
    int parts_left = 8;                  // wave-uniform: parts not yet found used up
    uint32_t err = 0, nsamples = 0, nnodes = 0;

    int mode = LANE_IDLE;
    uint32_t id = 0;
    Ray r;
    uint64_t node = 0;
    int nd = 0, k = 0, nprev = 0;
    uint8_t tent = 0;

    for (;;) {
        // ---- refill idle lanes
        const unsigned long long idle = __ballot(mode == LANE_IDLE);
        const int n_idle = __popcll(idle);
        if (parts_left > 0 && n_idle >= 16) {
            uint32_t base = 0, got = 0, beg = 0;
            // lane 0 claims items; the others take them by rank
            if (lane == 0) {
                while (parts_left > 0) {
                    uint32_t pb, pe;
                    part_range(total, unit, part, pb, pe);
                    uint32_t b0 = pe > pb ? atomicAdd(&a.next[part], (uint32_t)n_idle) : 0xffffffffu;
                    if (pe > pb && b0 < pe - pb) {
                        base = b0;
                        got = min((uint32_t)n_idle, pe - pb - b0);
                        beg = pb;
                        break;
                    }
                    part = (part + 1) & 7;
                    parts_left--;
                }
            }
            base = __shfl(base, 0);
            got = __shfl(got, 0);
            beg = __shfl(beg, 0);
            part = __shfl(part, 0);
            parts_left = __shfl(parts_left, 0);
            const uint32_t rank = (uint32_t)__popcll(idle & below);
            if (mode == LANE_IDLE && rank < got) {
                const uint32_t item = beg + base + rank;
                bool valid = true;
                if (first) {
                    const int t = (int)(item >> 8), q = (int)(item & 255u);
                    const int ty = t / a.tiles_x, tx = t % a.tiles_x;
                    const int w = q >> 6, ln = q & 63;
                    const int j = tx * 16 + (w & 1) * 8 + (ln & 7);
                    const int i = ty * 16 + (w >> 1) * 8 + (ln >> 3);
                    valid = i < H && j < W;
                    id = valid ? (uint32_t)(i * W + j) : 0u;
                } else {
                    id = a.in_list[item];
                }
                if (valid) {
                    make_ray(a.cam, (int)(id / (uint32_t)W), (int)(id % (uint32_t)W), r);
                    nd = 0;
                    k = 0;
                    if (first) {
                        node = root_ok(r, a.oct) ? next_node_seq(r, a.oct, 1ull, true) : 0ull;
                        nprev = 0;
                        tent = 0;
                        mode = LANE_WALK;
                    } else {
                        const uint8_t ph = a.ray_phase[id];
                        nprev = ph & PHASE_NMASK;
                        tent = ph & PHASE_TENT;
                        if (nprev > 0) {
                            mode = LANE_DEFER;
                        } else {
                            node = a.ray_node[id];
                            mode = LANE_WALK;
                        }
                    }
                }
            }
        }
        if (__ballot(mode != LANE_IDLE) == 0) {
            if (parts_left == 0) break;
            continue;
        }

        // ---- one node per active lane
        bool park = false;
        if (mode != LANE_IDLE) {
            const bool walk = mode == LANE_WALK;
            const uint64_t target = walk ? node : a.ray_defer[(size_t)k * R + id];
            const bool ended = walk && target == 0;
            uint64_t brick = 0, bidx = 0;
            int slot = -1;
            if (!ended) {
                brick = target >> bv.shift;
                bidx = brick - bv.brick_base;
                slot = bv.brick_slot[bidx];
            }
            float4 c;
            HitRecord h;
            bool hit = false;
            if (slot >= 0)
                hit = visit_node(r, a.oct, bv, sh, target, brick, slot, a.slot_ref, a.brick_trace, c, h, err, nsamples,
                                 nnodes);
            bool advance = false;
            if (walk) {
                if (ended || hit) {
                    if (ended) set_miss(c, h);
                    a.rgba[id] = c;
                    a.hits[id] = h;
                    if (nd > 0) {   // tentative: deferred nodes come first
                        a.ray_phase[id] = (uint8_t)(PHASE_TENT | nd);
                        park = true;
                    }
                    mode = LANE_IDLE;
                } else if (slot < 0) {
                    request_brick(a, bidx);
                    if (nd < a.max_defer) {
                        a.ray_defer[(size_t)nd * R + id] = target;
                        nd += 1;
                        advance = true;
                    } else {
                        a.ray_node[id] = target;
                        a.ray_phase[id] = (uint8_t)nd;
                        park = true;
                        mode = LANE_IDLE;
                    }
                } else {
                    advance = true;
                }
            } else {
                if (slot < 0) {   // still absent: keep the rest of the list, park again
                    for (int q = k; q < nprev; ++q)
                        a.ray_defer[(size_t)(q - k) * R + id] = a.ray_defer[(size_t)q * R + id];
                    request_brick(a, bidx);
                    a.ray_phase[id] = (uint8_t)(tent | (nprev - k));
                    park = true;
                    mode = LANE_IDLE;
                } else if (hit) {   // every earlier node missed: final
                    a.rgba[id] = c;
                    a.hits[id] = h;
                    mode = LANE_IDLE;
                } else if (++k == nprev) {
                    if (tent) {
                        mode = LANE_IDLE;   // the tentative result stands
                    } else {
                        node = a.ray_node[id];   // resume the walk
                        nd = 0;
                        mode = LANE_WALK;
                    }
                }
            }
            if (advance) node = next_node_seq(r, a.oct, target, false);
        }
        const uint32_t pslot = wave_atomic_add(&a.ctr->parked_count, park);
        if (park) a.out_list[pslot] = id;
    }

    if (err) atomicOr(&a.ctr->err, err);
    if (a.collect_stats) {
        unsigned long long s = nsamples, nn = nnodes;
        for (int off = 32; off > 0; off >>= 1) {
            s += __shfl_xor(s, off);
            nn += __shfl_xor(nn, off);
        }
        if (lane == 0) {
            atomicAdd(&a.ctr->samples, s);
            atomicAdd(&a.ctr->nodes, nn);
        }
    }
}

int render_persistent_workgroups(int device)
{
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        return 1024;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_render_persistent, kBlock, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    return cus * per_cu;
}

hipError_t launch_render_persistent(const RenderArgs& a, int max_workgroups, hipStream_t s)
{
    uint32_t total = a.in_list ? a.in_count : (uint32_t)a.tiles_total * 256u;
    if (total == 0) return hipSuccess;
    uint32_t grid = (total + kBlock - 1) / kBlock;
    if (grid > (uint32_t)max_workgroups) grid = (uint32_t)max_workgroups;
    hipLaunchKernelGGL(k_render_persistent, dim3(grid), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_render_pass(const RenderArgs& a, hipStream_t s)
{
    unsigned grid;
    if (a.in_list) {
        if (a.in_count == 0) return hipSuccess;
        grid = (a.in_count + kPassBlock - 1) / kPassBlock;
    } else {
        const unsigned tx8 = (unsigned)((a.cam.W + 7) / 8), ty8 = (unsigned)((a.cam.H + 7) / 8);
        if (a.supertile > 1) {
            // every XCD gets the same number of S x S blocks (the last ones may be empty)
            const unsigned S = (unsigned)a.supertile, blocks = ((tx8 + S - 1) / S) * ((ty8 + S - 1) / S);
            grid = (blocks + 7) / 8 * 8 * S * S;
        } else {
            const unsigned per = kPassBlock / 64;
            grid = (tx8 * ty8 + per - 1) / per;
        }
    }
    hipLaunchKernelGGL(k_render_pass, dim3(grid), dim3(kPassBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_render_wide(const RenderArgs& a, int lg, hipStream_t s)
{
    if (lg < 0 || lg > 6) return hipErrorInvalidValue;
    uint64_t rays;
    if (a.in_list) {
        rays = a.in_count;
    } else {
        rays = 64ull * (uint64_t)((a.cam.W + 7) / 8) * (uint64_t)((a.cam.H + 7) / 8);
    }
    if (rays == 0) return hipSuccess;
    // groups of 64 rays, padded to a multiple of 8 (one per XCD), 2^lg workgroups each
    const uint64_t groups = (rays + 63) / 64;
    const unsigned grid = (unsigned)(((groups + 7) / 8 * 8) << lg);
    hipLaunchKernelGGL(k_render_wide, dim3(grid), dim3(kPassBlock), 0, s, a, lg);
    return hipGetLastError();
}

hipError_t launch_render_multi(const RenderArgs& a, int lg, hipStream_t s)
{
    if (!a.jobs || a.nfresh < 0 || a.nfresh > kMaxFrames || lg < 0 || lg > 6) return hipErrorInvalidValue;
    const uint32_t tiles = a.nfresh > 0 ? a.fresh_tile0[a.nfresh] : 0u;
    const uint32_t list_groups = a.list_block0[kMaxFrames];
    if (tiles == 0 && lg > 0) {
        if (list_groups == 0) return hipSuccess;
        const uint64_t groups = list_groups;
        const unsigned grid = (unsigned)(((groups + 7) / 8 * 8) << lg);
        hipLaunchKernelGGL(k_render_multi_wide, dim3(grid), dim3(kPassBlock), 0, s, a, lg);
        return hipGetLastError();
    }
    const uint32_t list_region = a.supertile > 1 ? (list_groups + 7u) & ~7u : list_groups;
    if (a.supertile > 1 && tiles % 8u != 0) return hipErrorInvalidValue;   // frame regions in whole rounds of 8
    const uint64_t grid = (uint64_t)list_region + tiles;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_render_multi, dim3((unsigned)grid), dim3(kPassBlock), 0, s, a);
    return hipGetLastError();
}

// --------------------------------------------------------- serving pass
// The tiles of every admitted request (16x16 pixels, any size, scene or
// isovalue) in one launch.  A workgroup is one wave that renders an 8x8
// quarter of a tile (a group of 64 rays) with 2^lg lanes per ray (the wide
// march for light batches); the 2^lg waves of a group go to one XCD, as in
// k_render_wide.  Every brick of a serving scene is resident, so each ray
// finishes in the launch.  The grid has a fixed size per lg (a captured
// graph replays it); workgroups past the batch's groups exit.
template <bool kWide>
__device__ __forceinline__ void serve_ray(const ServeJob& J, int i, int j, const Lanes& g, uint32_t& err)
{
    Ray r;
    make_ray(J.cam, i, j, r);
    uint64_t node = root_ok(r, J.oct) ? next_node_seq(r, J.oct, 1ull, true) : 0ull;
    float4 c;
    HitRecord h;
    uint32_t ns = 0, nn = 0;
    for (;;) {
        if (node == 0) {
            set_miss(c, h);
            break;
        }
        const uint64_t brick = node >> J.bv.shift;
        const int slot = J.bv.brick_slot[brick - J.bv.brick_base];
        if (slot < 0) {   // cannot happen with a fully resident scene
            err |= 2u;
            set_miss(c, h);
            break;
        }
        if (visit_node_g<kWide>(r, J.oct, J.bv, J.sh, node, brick, slot, nullptr, nullptr, c, h, err, ns, nn, g))
            break;
        node = next_node_seq(r, J.oct, node, false);
    }
    if (g.leader) {
        const uint32_t id = (uint32_t)(i * J.cam.W + j);
        J.rgba[id] = c;
        J.hits[id] = h;
    }
}

template <bool kWide>
// at most 96 VGPRs: 5 waves per SIMD for both widths, without spilling
__global__ __launch_bounds__(kPassBlock, 5) void k_serve_pass(const ServeJob* __restrict__ jobs,
                                                           const ServeTile* __restrict__ work, ServeCounters* ctr,
                                                           int lg)
{
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t b = blockIdx.x;
    const uint32_t grp = ((b >> 3) >> lg) * 8u + (b & 7u);
    if (grp >= ctr->nwork * 4u) return;
    const uint32_t w = (b >> 3) & ((1u << lg) - 1u);
    const uint32_t p = (w << (6 - lg)) + ((uint32_t)lane >> lg);   // ray of the group, 0..63
    const ServeTile t = work[grp >> 2];
    const ServeJob& J = jobs[t.job];
    const int ty = (int)t.tile / J.tiles_x, tx = (int)t.tile % J.tiles_x;
    const int q = (int)(grp & 3u);
    const int i = ty * 16 + (q >> 1) * 8 + (int)(p >> 3);
    const int j = tx * 16 + (q & 1) * 8 + (int)(p & 7);
    if (i >= J.cam.H || j >= J.cam.W) return;   // the whole group
    const int L = 1 << lg;
    const Lanes g{L, lane & (L - 1), lane & ~(L - 1), (lane & (L - 1)) == 0};
    uint32_t err = 0;
    serve_ray<kWide>(J, i, j, g, err);
    if (err) atomicOr(&ctr->err, err);
}

hipError_t launch_serve_pass(const ServeJob* jobs, const ServeTile* work, ServeCounters* ctr, int max_tiles, int lg,
                             hipStream_t s)
{
    if (lg < 0 || lg > 6 || max_tiles <= 0) return hipErrorInvalidValue;
    const uint64_t groups = 4ull * (uint64_t)max_tiles;
    const unsigned grid = (unsigned)(((groups + 7) / 8 * 8) << lg);
    if (lg == 0)
        hipLaunchKernelGGL(k_serve_pass<false>, dim3(grid), dim3(kPassBlock), 0, s, jobs, work, ctr, lg);
    else
        hipLaunchKernelGGL(k_serve_pass<true>, dim3(grid), dim3(kPassBlock), 0, s, jobs, work, ctr, lg);
    return hipGetLastError();
}

// --------------------------------------------------------- traversal trace
// Emission sequence of the octree walk alone, as repeated calls of
// Octree::getBoxIntersected would produce it (Octree.cu:883-901).
__global__ __launch_bounds__(kBlock) void k_trace_nodes(TraceArgs a)
{
    uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    uint32_t R = (uint32_t)a.cam.W * (uint32_t)a.cam.H;
    if (id >= R) return;
    int i = (int)(id / (uint32_t)a.cam.W), j = (int)(id % (uint32_t)a.cam.W);
    Ray r;
    make_ray(a.cam, i, j, r);
    int count = 0;
    if (root_ok(r, a.oct)) {
        // the render passes' walker (ray-order children, the sorted list on ties)
        uint64_t node = next_node_seq(r, a.oct, 1ull, true);
        while (node != 0 && count < a.cap_total) {
            if (count < a.max_nodes) a.nodes[(size_t)id * a.max_nodes + count] = node;
            count++;
            node = next_node_seq(r, a.oct, node, false);
        }
    }
    a.counts[id] = count;
}

hipError_t launch_trace_nodes(const TraceArgs& a, hipStream_t s)
{
    uint32_t R = (uint32_t)a.cam.W * (uint32_t)a.cam.H;
    hipLaunchKernelGGL(k_trace_nodes, dim3((R + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------- staged protocol
// The reference's three per-pass stages on visibleCube_t arrays, for callers
// that drive Octree / lruCache / rayCaster themselves (visualTur.cu:175-220).

// Octree.cu:620-739 (cuda_getFirtsVoxel): next node of each NOCUBE ray.
// The resume point is the ray's current node code plus a 1-byte state in
// place of the reference's 32-entry stack.
__global__ __launch_bounds__(kBlock) void k_stage_traverse(StageTraverseArgs a)
{
    uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    uint32_t R = (uint32_t)a.cam.W * (uint32_t)a.cam.H;
    if (id >= R) return;
    Ray r;
    make_ray(a.cam, (int)(id / (uint32_t)a.cam.W), (int)(id % (uint32_t)a.cam.W), r);
    VisibleCube* c = a.cubes + id;
    if (!root_ok(r, a.oct)) {   // Octree.cu:729-733
        c->id = 0;
        return;
    }
    if (c->state != ST_NOCUBE) return;
    uint8_t ts = a.trav[id];
    if (ts == TRAV_EXHAUSTED) return;
    uint64_t node = (ts == TRAV_FRESH) ? next_node_seq(r, a.oct, 1ull, true) : next_node_seq(r, a.oct, c->id, false);
    if (node != 0) {
        c->id = node;
        c->state = ST_CUBE;
        a.trav[id] = TRAV_ACTIVE;
    } else {
        c->id = 0;
        a.trav[id] = TRAV_EXHAUSTED;
    }
}

hipError_t launch_stage_traverse(const StageTraverseArgs& a, hipStream_t s)
{
    uint32_t R = (uint32_t)a.cam.W * (uint32_t)a.cam.H;
    hipLaunchKernelGGL(k_stage_traverse, dim3((R + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

// rayCaster.cu:139-265 (cuda_rayCaster) on the staged state.
__global__ __launch_bounds__(kBlock) void k_stage_raycast(StageRaycastArgs a)
{
    uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    uint32_t R = (uint32_t)a.cam.W * (uint32_t)a.cam.H;
    if (id >= R) return;
    VisibleCube* c = a.cubes + id;
    if (c->state == ST_NOCUBE) {
        a.rgba[id] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        c->state = ST_PAINTED;
        return;
    }
    if (c->state != ST_CACHED) return;
    Ray r;
    make_ray(a.cam, (int)(id / (uint32_t)a.cam.W), (int)(id % (uint32_t)a.cam.W), r);
    int nx, ny, nz, bx, by, bz;
    node_min(c->id, a.level_o, a.n, nx, ny, nz);
    int side = 1 << (a.n - a.level_o);
    float tn, tf;
    if (!slab(r, nx, ny, nz, side, tn, tf)) return;   // state stays CACHED, as in the reference
    node_min(c->cubeID, a.level_c, a.n, bx, by, bz);
    bx -= a.inc; by -= a.inc; bz -= a.inc;
    uint32_t err = 0, ns = 0;
    float Xx, Xy, Xz;
    if (march_node(r, nx, ny, nz, side, c->data, a.S, bx, by, bz, a.iso, Xx, Xy, Xz, err, ns)) {
        float4 col;
        shade(c->data, a.S, bx, by, bz, Xx, Xy, Xz, a.lx, a.ly, a.lz, a.color_scale, col, err);
        a.rgba[id] = col;
        c->state = ST_PAINTED;
    } else {
        c->state = ST_NOCUBE;
    }
    if (err) atomicOr(a.err, err);
}

hipError_t launch_stage_raycast(const StageRaycastArgs& a, hipStream_t s)
{
    uint32_t R = (uint32_t)a.cam.W * (uint32_t)a.cam.H;
    hipLaunchKernelGGL(k_stage_raycast, dim3((R + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

// rayCaster.cu:267-357 (cuda_rayCasterCube): one brick, every pixel.  Its
// loop counter is a float and maxStep is not rounded up, and the colour
// ramp divides by the brick extent instead of 256.
__global__ __launch_bounds__(kBlock) void k_render_cube(CubeArgs a)
{
    uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    uint32_t R = (uint32_t)a.cam.W * (uint32_t)a.cam.H;
    if (id >= R) return;
    Ray r;
    make_ray(a.cam, (int)(id / (uint32_t)a.cam.W), (int)(id % (uint32_t)a.cam.W), r);
    uint32_t err = 0;
    bool hit = false;
    float Xx = 0.0f, Xy = 0.0f, Xz = 0.0f;
    const int S = a.dim + 2 * a.inc;
    const int bx = a.mnx - a.inc, by = a.mny - a.inc, bz = a.mnz - a.inc;
    float tnear, tfar;
    if (slab_box(r, a.mnx, a.mny, a.mnz, a.mnx + a.dim, a.mny + a.dim, a.mnz + a.dim, tnear, tfar)) {
        float X0 = r.ox + tnear * r.dx, X1 = r.oy + tnear * r.dy, X2 = r.oz + tnear * r.dz;
        float F0 = X0, F1 = X1, F2 = X2;
        float s0 = 0.5f * r.dx, s1 = 0.5f * r.dy, s2 = 0.5f * r.dz;
        float maxStep = (tfar - tnear) * 2.0f;
        const float cap = (float)(4 * a.dim + 8);
        if (maxStep > cap) maxStep = cap;
        bool first = true;
        float ant = 0.0f;
        for (float steps = 0.0f; steps <= maxStep; steps += 1.0f) {
            float v = sample(a.cube, S, bx, by, bz, X0, X1, X2, err);
            if (first) {
                first = false;
                ant = v;
                F0 = X0; F1 = X1; F2 = X2;
            } else if (((a.iso - ant) < 0.0f && (a.iso - v) < 0.0f) || ((a.iso - ant) > 0.0f && (a.iso - v) > 0.0f)) {
                ant = v;
                F0 = X0; F1 = X1; F2 = X2;
            } else {
                float t = (a.iso - ant) / (v - ant);
                Xx = F0 * (1.0f - t) + X0 * t;
                Xy = F1 * (1.0f - t) + X1 * t;
                Xz = F2 * (1.0f - t) + X2 * t;
                hit = true;
                break;
            }
            X0 = X0 + s0; X1 = X1 + s1; X2 = X2 + s2;
        }
    }
    if (hit) {
        float4 col;
        shade(a.cube, S, bx, by, bz, Xx, Xy, Xz, a.lx, a.ly, a.lz, (float)a.dim, col, err);
        a.rgba[id] = col;
    } else {
        a.rgba[id] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    }
    if (err) atomicOr(a.err, err);
}

hipError_t launch_render_cube(const CubeArgs& a, hipStream_t s)
{
    uint32_t R = (uint32_t)a.cam.W * (uint32_t)a.cam.H;
    hipLaunchKernelGGL(k_render_cube, dim3((R + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------ camera rays
__global__ void k_camera_rays(Camera cam, float* dirs)
{
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t R = (uint32_t)cam.W * (uint32_t)cam.H;
    if (id >= R) return;
    Ray r;
    make_ray(cam, (int)(id / (uint32_t)cam.W), (int)(id % (uint32_t)cam.W), r);
    dirs[3 * (size_t)id + 0] = r.dx;
    dirs[3 * (size_t)id + 1] = r.dy;
    dirs[3 * (size_t)id + 2] = r.dz;
}

hipError_t launch_camera_rays(const Camera& cam, float* dirs, hipStream_t s)
{
    uint32_t R = (uint32_t)cam.W * (uint32_t)cam.H;
    hipLaunchKernelGGL(k_camera_rays, dim3((R + 255) / 256), dim3(256), 0, s, cam, dirs);
    return hipGetLastError();
}

// ----------------------------------------------------------- brick staging
// One workgroup per brick: slot voxel (x,y,z) = global voxel
// (bmin - inc + (x,y,z)), zero outside the dataset (FileManager.cu:88-112).
__global__ __launch_bounds__(256) void k_bricks_from_volume(const float* __restrict__ vol,
                                                            int X, int Y, int Z,
                                                            const int64_t* __restrict__ codes,
                                                            const int32_t* __restrict__ slots,
                                                            int level_c, int n, int S, int inc, size_t stride,
                                                            float* __restrict__ pool)
{
    int b = blockIdx.x;
    int bx, by, bz;
    node_min((uint64_t)codes[b], level_c, n, bx, by, bz);
    bx -= inc; by -= inc; bz -= inc;
    float* dst = pool + (size_t)slots[b] * stride;
    int total = S * S * S;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        int z = e % S, y = (e / S) % S, x = e / (S * S);
        int gx = bx + x, gy = by + y, gz = bz + z;
        float v = 0.0f;
        if (gx >= 0 && gy >= 0 && gz >= 0 && gx < X && gy < Y && gz < Z)
            v = vol[((size_t)gx * Y + gy) * Z + gz];
        dst[e] = v;
    }
}

hipError_t launch_bricks_from_volume(const float* vol, int X, int Y, int Z,
                                     const int64_t* brick_codes, const int32_t* slots, int count,
                                     int level_c, int n, int cube_dim, int inc, size_t stride,
                                     float* pool, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    int S = cube_dim + 2 * inc;
    hipLaunchKernelGGL(k_bricks_from_volume, dim3(count), dim3(256), 0, s, vol, X, Y, Z,
                       brick_codes, slots, level_c, n, S, inc, stride, pool);
    return hipGetLastError();
}

// Whole-slot copies from device-visible pinned host memory.  Each brick is
// split over gridDim.x workgroups (blockIdx.x = part, blockIdx.y = brick;
// kCopySplit by default, fewer for copies that run beside a render pass)
// and every lane keeps four 16-byte reads in flight, so that enough PCIe
// reads are outstanding to cover their latency.  The slot size is a
// multiple of 4 floats when S is even; otherwise part 0 copies the tail.
constexpr int kCopySplit = 8;

__global__ __launch_bounds__(256) void k_copy_bricks(const float* const* __restrict__ src,
                                                     const int32_t* __restrict__ dst_slots,
                                                     size_t slot_elems, size_t stride, float* __restrict__ pool)
{
    const int b = blockIdx.y;
    const float* s = src[b];
    float* d = pool + (size_t)dst_slots[b] * stride;
    const size_t n4 = slot_elems / 4;
    const size_t per = (n4 + gridDim.x - 1) / gridDim.x;
    const size_t beg = per * blockIdx.x, end = beg + per < n4 ? beg + per : n4;
    const float4* __restrict__ s4 = reinterpret_cast<const float4*>(s);
    float4* __restrict__ d4 = reinterpret_cast<float4*>(d);
    size_t e = beg + threadIdx.x;
    for (; e + 768 < end; e += 1024) {
        float4 v0 = s4[e], v1 = s4[e + 256], v2 = s4[e + 512], v3 = s4[e + 768];
        d4[e] = v0;
        d4[e + 256] = v1;
        d4[e + 512] = v2;
        d4[e + 768] = v3;
    }
    for (; e < end; e += 256) d4[e] = s4[e];
    if (blockIdx.x == 0)
        for (size_t q = n4 * 4 + threadIdx.x; q < slot_elems; q += 256) d[q] = s[q];
}

hipError_t launch_copy_bricks(const float* const* src, const int32_t* dst_slots, int count,
                              size_t slot_elems, size_t stride, float* pool, hipStream_t s, int split)
{
    if (split <= 0) split = kCopySplit;
    for (int o = 0; o < count; o += 65535) {   // grid y is at most 65535
        const int m = count - o < 65535 ? count - o : 65535;
        hipLaunchKernelGGL(k_copy_bricks, dim3(split, m), dim3(256), 0, s, src + o, dst_slots + o, slot_elems,
                           stride, pool);
        if (hipError_t e = hipGetLastError()) return e;
    }
    return hipSuccess;
}

// Plain streaming copy: reference point for the HBM byte counters.
__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst,
                                                    size_t n4)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n4; e += (size_t)gridDim.x * blockDim.x)
        dst[e] = src[e];
}

hipError_t launch_copy_probe(const float* src, float* dst, size_t bytes, hipStream_t s)
{
    size_t n4 = bytes / 16;
    hipLaunchKernelGGL(k_copy_probe, dim3(2048), dim3(256), 0, s, reinterpret_cast<const float4*>(src),
                       reinterpret_cast<float4*>(dst), n4);
    return hipGetLastError();
}

// one workgroup per (slot, sub-block row): min and max of the 5^3 samples
// of each 4x4x4-cell sub-block of a brick's core
__global__ __launch_bounds__(64) void k_submm(const float* __restrict__ pool, const int32_t* __restrict__ slots,
                                              int S, int inc, int G, size_t stride, float2* __restrict__ submm)
{
    const int b = blockIdx.x;   // brick in the list
    const int slot = slots[b];
    const float* p = pool + (size_t)slot * stride;
    float2* out = submm + (size_t)slot * G * G * G;
    for (int sb = threadIdx.x; sb < G * G * G; sb += blockDim.x) {
        int gx = sb / (G * G), gy = (sb / G) % G, gz = sb % G;
        float lo = 3.0e38f, hi = -3.0e38f;
        bool nan = false;
        for (int x = 0; x <= 4; ++x)
            for (int y = 0; y <= 4; ++y)
                for (int z = 0; z <= 4; ++z) {
                    float v = p[((size_t)(inc + 4 * gx + x) * S + (inc + 4 * gy + y)) * S + (inc + 4 * gz + z)];
                    nan |= (v != v);
                    lo = fminf(lo, v);
                    hi = fmaxf(hi, v);
                }
        // a NaN sample makes every comparison false, so nothing is skipped
        out[sb] = nan ? make_float2(__builtin_nanf(""), __builtin_nanf("")) : make_float2(lo, hi);
    }
}

hipError_t launch_submm(const float* pool, const int32_t* slots, int count, int S, int inc, int G, size_t stride,
                        float2* submm, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_submm, dim3(count), dim3(64), 0, s, pool, slots, S, inc, G, stride, submm);
    return hipGetLastError();
}

// one thread per proven-side word: the sides of its 16 sub-blocks (fewer in
// a slot's last word), from their (min, max) as proven_side decides them
__global__ __launch_bounds__(256) void k_sides(const float2* __restrict__ submm, const int32_t* __restrict__ slots,
                                               int count, int G, float iso, uint32_t* __restrict__ sides,
                                               int side_words)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)count * side_words) return;
    const int b = (int)(i / (size_t)side_words), w = (int)(i % (size_t)side_words);
    const size_t slot = slots ? (size_t)slots[b] : (size_t)b;
    const int nsb = G * G * G;
    const float2* mm = submm + slot * nsb;
    uint32_t word = 0;
    for (int q = 0; q < 16; ++q) {
        const int sb = w * 16 + q;
        if (sb >= nsb) break;
        const float2 v = mm[sb];
        word |= (uint32_t)side_from_minmax(v.x, v.y, iso) << (2 * q);
    }
    sides[slot * side_words + w] = word;
}

hipError_t launch_sides(const float2* submm, const int32_t* slots, int count, int G, float iso, uint32_t* sides,
                        int side_words, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    const size_t n = (size_t)count * side_words;
    hipLaunchKernelGGL(k_sides, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, submm, slots, count, G, iso, sides,
                       side_words);
    return hipGetLastError();
}

__global__ void k_scatter_i32(int32_t* table, const int64_t* idx, const int32_t* val, int count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) table[idx[i]] = val[i];
}

hipError_t launch_scatter_i32(int32_t* table, const int64_t* idx, const int32_t* val, int count,
                              hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_i32, dim3((count + 255) / 256), dim3(256), 0, s, table, idx, val, count);
    return hipGetLastError();
}

__global__ void k_finish_service(int32_t* table, const int64_t* idx, const int32_t* val, int count, PassReset r)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = i; q < count; q += stride) table[idx[q]] = val[q];
    for (int64_t q = i; q < r.nwords; q += stride) r.req_bits[q] = 0u;
    for (int64_t q = i; q < r.nref; q += stride) r.ref[q] = 0;
    if (i < r.nctr) r.ctr[i] = 0u;
}

hipError_t launch_finish_service(int32_t* table, const int64_t* idx, const int32_t* val, int count,
                                 const PassReset& r, hipStream_t s)
{
    int64_t work = std::max<int64_t>(std::max<int64_t>(count, r.nwords), std::max<int64_t>(r.nref, r.nctr));
    if (work <= 0) return hipSuccess;
    int blocks = (int)std::min<int64_t>((work + 255) / 256, 1024);
    hipLaunchKernelGGL(k_finish_service, dim3(blocks), dim3(256), 0, s, table, idx, val, count, r);
    return hipGetLastError();
}

}  // namespace vt
