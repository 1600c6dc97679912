// vt_runtime.cpp -- host runtime behind the C ABI (include/visualtur_amd.h).
//
// Owns the device state of one rendering context: the octree membership
// tables, the brick pool with its page table and clock eviction, the pinned
// host tier filled from HDF5 by a reader thread, the per-ray parking state,
// and the pass loop of a frame.
#include "visualtur_amd.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "vt_build.hpp"
#include "vt_context.hpp"

using namespace vt;

namespace vt_rt {
thread_local std::string g_err;
int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}
}  // namespace vt_rt
using vt_rt::g_err;

namespace vt_rt {

int set_device(vt_context* ctx)
{
    HIP_TRY(hipSetDevice(ctx->device));
    return VT_OK;
}

OctreeView make_view(const vt_context* ctx, int final_level)
{
    OctreeView v;
    v.child_bytes = ctx->d_child;
    for (int l = 0; l < 24; ++l) v.level_off[l] = ctx->level_off[l];
    v.n = ctx->n;
    v.final_level = final_level;
    v.root_valid = ctx->root_valid;
    return v;
}

Camera to_cam(const vt_camera* c)
{
    Camera k;
    memcpy(&k, c, sizeof k);
    return k;
}

int check_camera(const vt_camera* cam)
{
    if (!cam) return fail(VT_E_ARG, "camera is NULL");
    if (cam->W <= 0 || cam->H <= 0) return fail(VT_E_ARG, "screen size must be positive");
    if ((int64_t)cam->W * cam->H > (1ll << 30)) return fail(VT_E_ARG, "screen too large");
    return VT_OK;
}

}  // namespace vt_rt

// Fold the device reference bits (set by the render kernel on every brick
// it marches) into the host clock.  With `copied`, the bytes were already
// read back into h_ref after the last pass, and the service clears them on
// the device; otherwise they are read and cleared here.
int fold_ref_bits(vt_context* ctx, bool copied)
{
    BrickCache& c = *ctx->cache;
    if (c.used == 0) return VT_OK;
    if (!copied) {
        HIP_TRY(hipMemcpyAsync(c.h_ref.p, c.d_ref, (size_t)c.gpu_slots, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    for (int64_t s = 0; s < c.gpu_slots; ++s) c.ref[(size_t)s] |= c.h_ref.p[(size_t)s];
    if (!copied) HIP_TRY(hipMemsetAsync(c.d_ref, 0, (size_t)c.gpu_slots, ctx->stream));
    return VT_OK;
}

int load_bricks(vt_context* ctx, const uint32_t* bricks, int64_t n_load, int64_t* misses, int32_t* slots_out,
                const int64_t* victims = nullptr, const PassReset* reset = nullptr);

// Starts the host tier once the brick geometry is known; with the octree
// uploaded down to levelC, a tier that holds every brick with octree nodes
// is filled ahead by the reader thread.
int host_tier_start(vt_context* ctx)
{
    BrickCache& c = *ctx->cache;
    if (!c.host || c.host->started) return VT_OK;
    const bool have = ctx->octree_ready && ctx->max_level >= c.level_c && (int)ctx->up_runs.size() > c.level_c;
    return c.host->start(ctx->n, have ? &ctx->up_runs[(size_t)c.level_c] : nullptr);
}

// Load `count` requested bricks (indices relative to 8^levelC) into the
// device pool as far as capacity allows.  Bricks loaded here are pinned
// against eviction until the next call.  Between two render passes
// (`reset` given) the reference bytes were read back with the pass counters,
// the work is only enqueued (the next pass follows on the same stream), and
// the last kernel of the service also resets the per-pass state.
int service_requests(vt_context* ctx, const uint32_t* reqs, int64_t count, vt_frame_stats* st,
                     const PassReset* reset)
{
    BrickCache& c = *ctx->cache;
    if (int rc = fold_ref_bits(ctx, reset != nullptr)) return rc;
    std::vector<uint32_t> sorted;   // bricks still absent (a load on the copy stream may have placed some)
    sorted.reserve((size_t)count);
    for (int64_t i = 0; i < count; ++i)
        if (c.table[(size_t)reqs[i]] < 0) sorted.push_back(reqs[i]);
    std::sort(sorted.begin(), sorted.end());
    sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
    int64_t n_load = std::min<int64_t>((int64_t)sorted.size(), c.gpu_slots);
    if (c.host && !c.host->full_mode) n_load = std::min<int64_t>(n_load, c.host->capacity);
    c.epoch++;
    int64_t misses = 0;
    if (int rc = load_bricks(ctx, sorted.data(), n_load, &misses, nullptr, nullptr, reset)) return rc;
    if (n_load <= 0 && reset) HIP_TRY(launch_finish_service(c.d_table, nullptr, nullptr, 0, *reset, ctx->stream));
    c.stat_missed += count;
    if (st) {
        st->bricks_uploaded += n_load;
        st->host_tier_misses += misses;
    }
    return VT_OK;
}

// Place `n_load` non-resident bricks into victim slots (skipping slots
// pinned in the current epoch), then fill them from the brick source.  The
// staging arrays go to the device in one copy.  Without `reset` the call
// waits for the loads; with it (between two render passes) the loads are
// only enqueued, their last kernel also resets the per-pass state, and the
// staging is next rewritten after the following pass has been waited for.
int load_bricks(vt_context* ctx, const uint32_t* bricks, int64_t n_load, int64_t* misses, int32_t* slots_out,
                const int64_t* victims, const PassReset* reset)
{
    BrickCache& c = *ctx->cache;
    if (n_load <= 0) return VT_OK;
    if (int rc = copy_stream_idle(ctx)) return rc;   // evictions below must not hit a slot still being filled
    const StageLayout L((size_t)n_load);
    if (c.h_stage.ensure(L.bytes) != hipSuccess || c.d_stage.ensure(L.bytes) != hipSuccess)
        return fail(VT_E_NOMEM, "staging allocation failed");
    const float** h_src = reinterpret_cast<const float**>(c.h_stage.p + L.src);
    int64_t* h_codes = reinterpret_cast<int64_t*>(c.h_stage.p + L.codes);
    int64_t* h_tidx = reinterpret_cast<int64_t*>(c.h_stage.p + L.tidx);
    int32_t* h_slots = reinterpret_cast<int32_t*>(c.h_stage.p + L.slots);
    int32_t* h_tval = reinterpret_cast<int32_t*>(c.h_stage.p + L.tval);
    int64_t nt = 0;
    int fail_rc = VT_OK;   // a brick source that fails: the bricks before it still load
    for (int64_t q = 0; q < n_load; ++q) {
        int64_t b = bricks[q];
        int64_t slot = victims ? victims[q] : c.pick_victim();
        if (slot < 0) return fail(VT_E_STATE, "brick cache: no evictable slot");
        int64_t old = c.slot_brick[(size_t)slot];
        if (old >= 0) {
            c.table[(size_t)old] = -1;
            h_tidx[nt] = old; h_tval[nt] = -1; nt++;
            c.stat_evictions++;
        }
        c.slot_brick[(size_t)slot] = b;
        c.table[(size_t)b] = (int32_t)slot;
        h_tidx[nt] = b; h_tval[nt] = (int32_t)slot; nt++;
        c.pin[(size_t)slot] = c.epoch;
        c.ref[(size_t)slot] = 1;
        h_slots[q] = (int32_t)slot;
        h_codes[q] = b + (int64_t)(1ull << (3 * c.level_c));
        if (slots_out) slots_out[q] = (int32_t)slot;
        h_src[q] = nullptr;
        if (c.host) {
            const float* src = nullptr;
            int rc = c.host->get(b, c.epoch, &src, misses);
            if (rc != VT_OK) {
                // b stays absent and its slot free (an evicted brick stays
                // evicted); the bricks staged before it are loaded below, so
                // the host and device tables agree when the error returns
                c.table[(size_t)b] = -1;
                c.slot_brick[(size_t)slot] = -1;
                c.ref[(size_t)slot] = 0;
                c.free_list.push_back(slot);
                nt--;
                fail_rc = rc;
                n_load = q;
                break;
            }
            h_src[q] = src;
        }
    }
    // everything up to the last used table value, in one copy
    HIP_TRY(hipMemcpyAsync(c.d_stage.p, c.h_stage.p, L.tval + (size_t)nt * 4, hipMemcpyHostToDevice, ctx->stream));
    const float* const* d_src = reinterpret_cast<const float* const*>(c.d_stage.p + L.src);
    const int64_t* d_codes = reinterpret_cast<const int64_t*>(c.d_stage.p + L.codes);
    const int64_t* d_tidx = reinterpret_cast<const int64_t*>(c.d_stage.p + L.tidx);
    const int32_t* d_slots = reinterpret_cast<const int32_t*>(c.d_stage.p + L.slots);
    const int32_t* d_tval = reinterpret_cast<const int32_t*>(c.d_stage.p + L.tval);
    if (c.host) {
        HIP_TRY(launch_copy_bricks(d_src, d_slots, (int)n_load, c.brick_elems, c.slot_elems, c.d_pool, ctx->stream));
    } else {
        HIP_TRY(launch_bricks_from_volume(c.d_vol, c.vol_dims[0], c.vol_dims[1], c.vol_dims[2], d_codes, d_slots,
                                          (int)n_load, c.level_c, c.n, c.cube, c.inc, c.slot_elems, c.d_pool,
                                          ctx->stream));
    }
    if (c.d_submm)
        HIP_TRY(launch_submm(c.d_pool, d_slots, (int)n_load, c.S, c.inc, c.G, c.slot_elems, c.d_submm, ctx->stream));
    if (c.d_sides && c.side_iso == c.side_iso)
        HIP_TRY(launch_sides(c.d_submm, d_slots, (int)n_load, c.G, c.side_iso, c.d_sides, c.side_words, ctx->stream));
    if (reset) {
        HIP_TRY(launch_finish_service(c.d_table, d_tidx, d_tval, (int)nt, *reset, ctx->stream));
    } else {
        HIP_TRY(launch_scatter_i32(c.d_table, d_tidx, d_tval, (int)nt, ctx->stream));
        // the staging buffers are reused by the next call: finish the copies
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    c.stat_uploads += n_load;
    return fail_rc;
}

// Loads the absent bricks of `req` that fit in free slots on the copy
// stream, without waiting; the others are appended to `left`.  With
// `publish`, the sub-block min/max and page-table entries follow on the copy
// stream (no render pass may run then: vt_render_frame waits for the copy
// stream before its first pass); without, the copies go alone (a render pass
// may run beside them, it reads no free slot) and publish_loads writes the
// rest later, between passes.  Host-tier slots read here stay pinned in the
// current epoch, which ends only after the copy stream has been synchronized.
int load_async(vt_context* ctx, const uint32_t* req, size_t n, std::vector<uint32_t>& left, int64_t* loaded,
               int64_t* misses, bool publish)
{
    BrickCache& c = *ctx->cache;
    AsyncStage& A = ctx->astage;
    const int64_t base = A.used;
    int64_t k = 0;
    bool full = false;
    int fail_rc = VT_OK;   // a brick source that fails: the bricks before it still load
    for (size_t i = 0; i < n; ++i) {
        const uint32_t b = req[i];
        if (c.table[(size_t)b] >= 0) continue;
        int64_t slot = full || base + k >= A.cap ? -1 : c.take_free();
        if (slot < 0) {
            full = true;
            left.push_back(b);
            continue;
        }
        const float* src = nullptr;
        if (c.host) {
            int rc = c.host->get(b, c.epoch, &src, misses);
            if (rc == VT_E_STATE) {   // every host slot pinned in this epoch: leave it to the synchronous service
                c.free_list.push_back(slot);
                full = true;
                left.push_back(b);
                continue;
            }
            if (rc != VT_OK) {
                c.free_list.push_back(slot);
                fail_rc = rc;
                break;
            }
        }
        c.slot_brick[(size_t)slot] = b;
        c.table[(size_t)b] = (int32_t)slot;
        c.pin[(size_t)slot] = c.epoch;
        c.ref[(size_t)slot] = 1;
        A.src.p[base + k] = src;
        A.slots.p[base + k] = (int32_t)slot;
        A.codes.p[base + k] = (int64_t)b + (int64_t)(1ull << (3 * c.level_c));
        A.tidx.p[base + k] = b;
        A.tval.p[base + k] = (int32_t)slot;
        k++;
    }
    if (k == 0) return fail_rc;
    hipStream_t s = ctx->cstream;
    if (c.host) {
        // beside a running render pass, two workgroups per brick keep enough
        // PCIe reads in flight while taking few of the pass's wave slots
        HIP_TRY(launch_copy_bricks(A.src.dev + base, A.slots.dev + base, (int)k, c.brick_elems, c.slot_elems,
                                   c.d_pool, s, publish ? 0 : 2));
    } else {
        HIP_TRY(launch_bricks_from_volume(c.d_vol, c.vol_dims[0], c.vol_dims[1], c.vol_dims[2], A.codes.dev + base,
                                          A.slots.dev + base, (int)k, c.level_c, c.n, c.cube, c.inc, c.slot_elems, c.d_pool, s));
    }
    A.used += k;
    if (publish) {
        if (int rc = publish_loads(ctx, s)) return rc;
    }
    c.stat_uploads += k;
    *loaded += k;
    return fail_rc;
}

// Writes the sub-block min/max and the page-table entries of the copies not
// yet published, on stream `s` after the copy stream's work so far.  Only
// called where no render pass runs: a running pass never sees a page-table
// entry change (an entry written beside it could expose a slot whose lines
// another XCD's L2 still holds from before the copy).
int publish_loads(vt_context* ctx, hipStream_t s)
{
    BrickCache& c = *ctx->cache;
    AsyncStage& A = ctx->astage;
    const int64_t k = A.used - A.pub;
    if (k <= 0) return VT_OK;
    if (s != ctx->cstream) {
        HIP_TRY(hipEventRecord(ctx->ev_copy, ctx->cstream));
        HIP_TRY(hipStreamWaitEvent(s, ctx->ev_copy, 0));
    }
    if (c.d_submm)
        HIP_TRY(launch_submm(c.d_pool, A.slots.dev + A.pub, (int)k, c.S, c.inc, c.G, c.slot_elems, c.d_submm, s));
    if (c.d_sides && c.side_iso == c.side_iso)
        HIP_TRY(launch_sides(c.d_submm, A.slots.dev + A.pub, (int)k, c.G, c.side_iso, c.d_sides, c.side_words, s));
    HIP_TRY(launch_scatter_i32(c.d_table, A.tidx.dev + A.pub, A.tval.dev + A.pub, (int)k, s));
    A.pub = A.used;
    return VT_OK;
}

// Waits for the copy stream (publishing what it loaded) and for the
// publication; its staging can then be reused.
int copy_stream_idle(vt_context* ctx)
{
    if (ctx->astage.pub < ctx->astage.used) {
        if (int rc = publish_loads(ctx, ctx->stream)) return rc;
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->cstream));
    ctx->astage.used = ctx->astage.pub = 0;
    return VT_OK;
}

// Makes the proven-side words hold the sides for `iso` (a no-op when they
// already do, or for a NaN iso): recomputed for every slot on the context
// stream, where the sub-block (min, max) of every published load is already
// written; the copy stream's later work waits for them.  A ray whose iso is
// not side_iso reads the (min, max) instead.
int ensure_sides(vt_context* ctx, float iso)
{
    BrickCache& c = *ctx->cache;
    if (!c.d_sides || iso != iso || iso == c.side_iso) return VT_OK;
    c.side_iso = iso;
    HIP_TRY(launch_sides(c.d_submm, nullptr, (int)c.gpu_slots, c.G, iso, c.d_sides, c.side_words, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev_copy, ctx->stream));
    HIP_TRY(hipStreamWaitEvent(ctx->cstream, ctx->ev_copy, 0));
    return VT_OK;
}

// Frees slots (clock victims, page-table entries cleared on the context
// stream) until `reserve` slots are free, for the loads issued while the
// first pass runs.
int make_reserve(vt_context* ctx, int64_t reserve)
{
    BrickCache& c = *ctx->cache;
    int64_t need = std::min<int64_t>(reserve, c.gpu_slots) - c.available();
    if (need <= 0) return VT_OK;
    if (int rc = fold_ref_bits(ctx)) return rc;
    const StageLayout L((size_t)need);
    if (c.h_stage.ensure(L.bytes) != hipSuccess || c.d_stage.ensure(L.bytes) != hipSuccess)
        return fail(VT_E_NOMEM, "staging allocation failed");
    int64_t* h_tidx = reinterpret_cast<int64_t*>(c.h_stage.p + L.tidx);
    int32_t* h_tval = reinterpret_cast<int32_t*>(c.h_stage.p + L.tval);
    int64_t nt = 0;
    for (; nt < need; ++nt) {
        int64_t slot = c.clock_victim(true);
        if (slot < 0) slot = c.clock_victim(false);
        if (slot < 0) break;
        const int64_t old = c.slot_brick[(size_t)slot];
        c.table[(size_t)old] = -1;
        c.slot_brick[(size_t)slot] = -1;
        c.free_list.push_back(slot);
        c.stat_evictions++;
        h_tidx[nt] = old;
        h_tval[nt] = -1;
    }
    if (nt == 0) return VT_OK;
    HIP_TRY(hipMemcpyAsync(c.d_stage.p + L.tidx, c.h_stage.p + L.tidx, (size_t)nt * 8, hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(c.d_stage.p + L.tval, c.h_stage.p + L.tval, (size_t)nt * 4, hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(launch_scatter_i32(c.d_table, reinterpret_cast<const int64_t*>(c.d_stage.p + L.tidx),
                               reinterpret_cast<const int32_t*>(c.d_stage.p + L.tval), (int)nt, ctx->stream));
    return VT_OK;
}

// While the first pass runs (until `done` completes), loads the bricks it
// requests into the reserved free slots on the copy stream, without
// publishing them; requests that find no free slot are left to the service
// after the pass.
int load_during_pass(vt_context* ctx, hipEvent_t done, int64_t* loaded, int64_t* misses)
{
    BrickCache& c = *ctx->cache;
    const uint32_t* rq = ctx->req_list.p;
    const size_t cap = (size_t)c.nbricks;
    std::vector<uint32_t> batch, left;
    size_t seen = 0;
    double last = now_ms();
    for (;;) {
        const hipError_t q = hipEventQuery(done);
        if (q != hipSuccess && q != hipErrorNotReady) HIP_TRY(q);
        if (q == hipSuccess) return VT_OK;   // the service after the pass loads the rest
        while (seen < cap) {
            const uint32_t v = __atomic_load_n(&rq[seen], __ATOMIC_ACQUIRE);
            if (v == 0) break;
            batch.push_back(v - 1);
            ++seen;
        }
        const double t = now_ms();
        if (!batch.empty() && (batch.size() >= 16 || t - last > 0.05)) {
            left.clear();
            if (int rc = load_async(ctx, batch.data(), batch.size(), left, loaded, misses, false)) return rc;
            batch.clear();
            last = t;
            if (!left.empty()) return VT_OK;   // the reserve is used up
        }
    }
}

namespace {

bool runs_well_formed(const uint64_t* E, int size, int level, std::string& why)
{
    if (size < 0 || (size & 1)) { why = "odd run-list size"; return false; }
    uint64_t lo = 1ull << (3 * level), hi = (2ull << (3 * level)) - 1;
    for (int j = 0; j < size; j += 2) {
        if (E[j] > E[j + 1]) { why = "run start after its end"; return false; }
        if (E[j] < lo || E[j + 1] > hi) { why = "code outside its level"; return false; }
        if (j >= 2 && E[j] <= E[j - 1]) { why = "runs not sorted and disjoint"; return false; }
    }
    return true;
}

}  // namespace

// ===================================================================== ABI

// Per-ray state of the render passes: resume node, deferral lists
// ([max_defer][R], pass-major), the parked-ray lists of consecutive passes,
// and the phase byte.
struct RayState {
    uint64_t* node;
    uint64_t* defer;
    uint32_t* list[2];
    uint8_t* phase;
};

// ------------------------------------------------- allocation-free frames
// vt_render_frame_noalloc keeps the per-ray state of its render passes in
// this array instead of buffers the context allocates, for frames of at most
// kScratchRays rays (1920x1080 fits) with deferral lists of at most
// kScratchDefer entries: 33 B per ray, 66 MiB.  Every GPU the library runs
// on holds its own copy until the process ends.  Host code never passes the
// array's name to a kernel: it uses the device address looked up once per
// device at context creation (scratch_address).
constexpr uint32_t kScratchRays = 1u << 21;
constexpr int kScratchDefer = 2;
constexpr size_t kScratchBytes = (size_t)kScratchRays * (8 + 8 * kScratchDefer + 4 + 4 + 1);
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_m;
uint8_t* g_scratch[kMaxDevices] = {};    // device address of `scratch` on each device
std::mutex g_scratch_use[kMaxDevices];   // one allocation-free frame at a time per device

// Looks up the current device's copy of `scratch` (once per device).
int scratch_address(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(VT_E_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_scratch_m);
    if (g_scratch[device]) return VT_OK;
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(VT_E_HIP, "hipGetSymbolAddress returned no address for the frame scratch");
    g_scratch[device] = static_cast<uint8_t*>(ptr);
    return VT_OK;
}
}  // namespace

extern "C" {

int vt_last_error(vt_context*, char* buf, size_t len)
{
    if (!buf || len == 0) return VT_E_ARG;
    snprintf(buf, len, "%s", g_err.c_str());
    return VT_OK;
}

int vt_context_create(int device, vt_context** out)
{
    if (!out) return fail(VT_E_ARG, "out is NULL");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0) HIP_TRY(hipGetDevice(&device));   // -1: the calling thread's current device
    if (device < 0 || device >= ndev) return fail(VT_E_ARG, "no such device");
    std::unique_ptr<vt_context> ctx(new vt_context);
    ctx->device = device;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&ctx->cstream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&ctx->ev0));
    HIP_TRY(hipEventCreate(&ctx->ev1));
    HIP_TRY(hipEventCreateWithFlags(&ctx->ev_copy, hipEventDisableTiming));
    HIP_TRY(hipMalloc((void**)&ctx->d_ctr, sizeof(FrameCounters)));
    if (ctx->d_next.ensure(8) != hipSuccess) return fail(VT_E_NOMEM, "work counters");
    ctx->persistent_wgs = render_persistent_workgroups(device);
    if (int rc = scratch_address(device)) return rc;
    HIP_TRY(hipHostMalloc((void**)&ctx->h_ctr, sizeof(FrameCounters), hipHostMallocDefault));
    *out = ctx.release();
    return VT_OK;
}

int vt_context_destroy(vt_context* ctx)
{
    if (!ctx) return VT_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return VT_OK;
}

int vt_synchronize(vt_context* ctx)
{
    if (!ctx) return fail(VT_E_ARG, "ctx is NULL");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VT_OK;
}

int vt_device_memory(int32_t device, size_t* free_bytes, size_t* total_bytes)
{
    if (!free_bytes || !total_bytes) return fail(VT_E_ARG, "NULL argument");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(VT_E_ARG, "no such device");
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipMemGetInfo(free_bytes, total_bytes));
    return VT_OK;
}

int vt_malloc(vt_context* ctx, size_t bytes, void** out)
{
    if (!ctx || !out) return fail(VT_E_ARG, "NULL argument");
    if (int rc = set_device(ctx)) return rc;
    HIP_TRY(hipMalloc(out, std::max<size_t>(bytes, 1)));
    return VT_OK;
}

int vt_free(vt_context* ctx, void* p)
{
    if (!ctx) return fail(VT_E_ARG, "ctx is NULL");
    if (!p) return VT_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipFree(p));
    return VT_OK;
}

int vt_memcpy_h2d(vt_context* ctx, void* d_dst, const void* src, size_t bytes)
{
    if (!ctx || (bytes && (!d_dst || !src))) return fail(VT_E_ARG, "NULL argument");
    HIP_TRY(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VT_OK;
}

int vt_memcpy_d2h(vt_context* ctx, void* dst, const void* d_src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !d_src))) return fail(VT_E_ARG, "NULL argument");
    HIP_TRY(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VT_OK;
}

int vt_memset(vt_context* ctx, void* d_dst, int value, size_t bytes)
{
    if (!ctx || (bytes && !d_dst)) return fail(VT_E_ARG, "NULL argument");
    HIP_TRY(hipMemsetAsync(d_dst, value, bytes, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VT_OK;
}

int vt_screen_size(float distance, float fovH, float fovW, float* h, float* w)
{
    if (!h || !w) return fail(VT_E_ARG, "NULL argument");
    // Screen.cpp:18-19
    *h = 2 * distance * tanf((float)(fovH * (M_PI / 180.0)));
    *w = 2 * distance * tanf((float)(fovW * (M_PI / 180.0)));
    return VT_OK;
}

int vt_octree_upload(vt_context* ctx, const uint64_t* runs, const int32_t* sizes, int32_t nLevels,
                     int32_t maxLevel)
{
    if (!ctx || !runs || !sizes) return fail(VT_E_ARG, "NULL argument");
    if (nLevels < 1 || nLevels > 20) return fail(VT_E_ARG, "nLevels must be in [1,20]");
    if (maxLevel < 0 || maxLevel > nLevels || maxLevel > 9)
        return fail(VT_E_ARG, "maxLevel must be in [0, min(nLevels, 9)]");
    if (int rc = set_device(ctx)) return rc;
    // validate and expand the run lists into one membership byte per parent
    std::vector<uint64_t> off(24, 0);
    uint64_t total = 0;
    for (int l = 1; l <= maxLevel; ++l) {
        off[l] = total;
        total += 1ull << (3 * (l - 1));
    }
    std::vector<uint8_t> bytes((size_t)std::max<uint64_t>(total, 1), 0);
    const uint64_t* E = runs;
    int root = 0;
    for (int l = 0; l <= maxLevel; ++l) {
        std::string why;
        if (!runs_well_formed(E, sizes[l], l, why))
            return fail(VT_E_FORMAT, "octree level " + std::to_string(l) + ": " + why);
        if (l == 0) {
            root = sizes[0] >= 2 && E[0] <= 1 && 1 <= E[1];
        } else {
            uint64_t base = 1ull << (3 * l);
            for (int j = 0; j < sizes[l]; j += 2)
                for (uint64_t c = E[j]; c <= E[j + 1]; ++c) {
                    uint64_t parent_idx = (c >> 3) - (base >> 3);
                    bytes[(size_t)(off[l] + parent_idx)] |= (uint8_t)(1u << (c & 7));
                }
        }
        E += sizes[l];
    }
    ctx->up_runs.assign((size_t)maxLevel + 1, {});
    {
        const uint64_t* R = runs;
        for (int l = 0; l <= maxLevel; ++l) {
            ctx->up_runs[(size_t)l].assign(R, R + sizes[l]);
            R += sizes[l];
        }
    }
    ctx->cand_level = -1;
    if (ctx->d_child) (void)hipFree(ctx->d_child);
    ctx->d_child = nullptr;
    HIP_TRY(hipMalloc((void**)&ctx->d_child, bytes.size()));
    HIP_TRY(hipMemcpy(ctx->d_child, bytes.data(), bytes.size(), hipMemcpyHostToDevice));
    for (int l = 0; l < 24; ++l) ctx->level_off[l] = off[l];
    ctx->n = nLevels;
    ctx->max_level = maxLevel;
    ctx->root_valid = root;
    ctx->octree_ready = true;
    return VT_OK;
}

int vt_octree_load_file(vt_context* ctx, const char* path, int32_t maxLevel, float* iso, int32_t* nLevels,
                        int32_t realDim[3])
{
    if (!ctx || !path) return fail(VT_E_ARG, "NULL argument");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(VT_E_IO, std::string("cannot open octree file ") + path);
    int32_t magic = 0, dim = 0, rd[3] = {0, 0, 0}, n = 0;
    float is = 0.0f;
    bool ok = fread(&magic, 4, 1, f) == 1;
    if (!ok || magic != 919278872) {   // Octree.cu:781
        fclose(f);
        return fail(VT_E_FORMAT, "invalid octree file format (magic)");
    }
    ok = fread(&is, 4, 1, f) == 1 && fread(&dim, 4, 1, f) == 1 && fread(rd, 4, 3, f) == 3 &&
         fread(&n, 4, 1, f) == 1;
    if (!ok || n < 1 || n > 20 || maxLevel > n) {
        fclose(f);
        return fail(VT_E_FORMAT, "truncated or inconsistent octree header");
    }
    std::vector<std::vector<uint64_t>> lv((size_t)n + 1);
    for (int l = n; l >= 0; --l) {
        int32_t ne = 0;
        if (fread(&ne, 4, 1, f) != 1 || ne < 0) { fclose(f); return fail(VT_E_FORMAT, "truncated octree level"); }
        lv[(size_t)l].resize((size_t)ne);
        if (ne && fread(lv[(size_t)l].data(), 8, (size_t)ne, f) != (size_t)ne) {
            fclose(f);
            return fail(VT_E_FORMAT, "truncated octree level");
        }
    }
    fclose(f);
    std::vector<uint64_t> flat;
    std::vector<int32_t> sizes((size_t)maxLevel + 1);
    for (int l = 0; l <= maxLevel; ++l) {
        sizes[(size_t)l] = (int32_t)lv[(size_t)l].size();
        flat.insert(flat.end(), lv[(size_t)l].begin(), lv[(size_t)l].end());
    }
    int rc = vt_octree_upload(ctx, flat.data(), sizes.data(), n, maxLevel);
    if (rc != VT_OK) return rc;
    if (iso) *iso = is;
    if (nLevels) *nLevels = n;
    if (realDim) { realDim[0] = rd[0]; realDim[1] = rd[1]; realDim[2] = rd[2]; }
    return VT_OK;
}

int vt_camera_rays(vt_context* ctx, const vt_camera* cam, float* d_dirs)
{
    if (!ctx || !d_dirs) return fail(VT_E_ARG, "NULL argument");
    if (int rc = check_camera(cam)) return rc;
    if (int rc = set_device(ctx)) return rc;
    HIP_TRY(launch_camera_rays(to_cam(cam), d_dirs, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VT_OK;
}

int vt_cache_create(vt_context* ctx, int32_t levelC, int32_t cubeDim, int32_t cubeInc, int64_t gpuSlots,
                    int64_t cpuSlots)
{
    if (!ctx) return fail(VT_E_ARG, "ctx is NULL");
    if (levelC < 0 || levelC > 9) return fail(VT_E_ARG, "levelC must be in [0,9]");
    if (cubeDim <= 0 || cubeInc < 0 || gpuSlots <= 0 || cpuSlots < 0)
        return fail(VT_E_ARG, "bad cache dimensions");
    int64_t S = cubeDim + 2 * (int64_t)cubeInc;
    if (S > 1024) return fail(VT_E_ARG, "brick too large");
    ctx->early_last = INT64_MAX;
    ctx->pass_state_clean = false;   // the request bitmap may be reallocated
    if (int rc = set_device(ctx)) return rc;
    std::unique_ptr<BrickCache> c(new BrickCache);
    c->level_c = levelC;
    c->cube = cubeDim;
    c->inc = cubeInc;
    c->S = (int)S;
    c->brick_elems = (size_t)(S * S * S);
    c->slot_elems = (c->brick_elems + 63) / 64 * 64;
    c->nbricks = 1ll << (3 * levelC);
    c->gpu_slots = std::min<int64_t>(gpuSlots, c->nbricks);
    c->cpu_slots = cpuSlots;
    size_t pool_bytes = (size_t)c->gpu_slots * c->slot_elems * sizeof(float);
    if (hipMalloc((void**)&c->d_pool, pool_bytes) != hipSuccess)
        return fail(VT_E_NOMEM, "brick pool allocation failed (" + std::to_string(pool_bytes) + " bytes)");
    HIP_TRY(hipMalloc((void**)&c->d_table, (size_t)c->nbricks * 4));
    HIP_TRY(hipMemset(c->d_table, 0xff, (size_t)c->nbricks * 4));
    HIP_TRY(hipMalloc((void**)&c->d_ref, (size_t)c->gpu_slots));
    HIP_TRY(hipMemset(c->d_ref, 0, (size_t)c->gpu_slots));
    if (cubeDim % 4 == 0 && cubeInc >= 1) {   // sub-blocks need the +1 sample of the halo
        c->G = cubeDim / 4;
        size_t mm = (size_t)c->gpu_slots * c->G * c->G * c->G * sizeof(float2);
        if (hipMalloc((void**)&c->d_submm, mm) != hipSuccess) return fail(VT_E_NOMEM, "sub-block min/max");
        c->side_words = (c->G * c->G * c->G + 16 * 32 - 1) / (16 * 32) * 32;
        if (hipMalloc((void**)&c->d_sides, (size_t)c->gpu_slots * c->side_words * 4) != hipSuccess)
            return fail(VT_E_NOMEM, "sub-block sides");
    }
    c->table.assign((size_t)c->nbricks, -1);
    c->slot_brick.assign((size_t)c->gpu_slots, -1);
    c->ref.assign((size_t)c->gpu_slots, 0);
    if (c->h_ref.ensure((size_t)c->gpu_slots) != hipSuccess) return fail(VT_E_NOMEM, "pinned reference bytes");
    c->pin.assign((size_t)c->gpu_slots, 0);
    ctx->cache = std::move(c);
    // request structures sized to the brick count
    if (ctx->req_bits.ensure((size_t)((ctx->cache->nbricks + 31) / 32)) != hipSuccess ||
        ctx->req_list.ensure((size_t)ctx->cache->nbricks, hipHostMallocCoherent) != hipSuccess ||
        ctx->astage.ensure(ctx->cache->gpu_slots) != hipSuccess)
        return fail(VT_E_NOMEM, "request buffers");
    // staging of the largest load (every slot), so that no frame allocates it
    const StageLayout L((size_t)ctx->cache->gpu_slots);
    if (ctx->cache->h_stage.ensure(L.bytes) != hipSuccess || ctx->cache->d_stage.ensure(L.bytes) != hipSuccess)
        return fail(VT_E_NOMEM, "staging allocation failed");
    memset(ctx->req_list.p, 0, (size_t)ctx->cache->nbricks * 4);
    return VT_OK;
}

int vt_cache_destroy(vt_context* ctx)
{
    if (!ctx) return fail(VT_E_ARG, "ctx is NULL");
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->cstream);
    ctx->cache.reset();
    return VT_OK;
}

int vt_cache_attach_device_volume(vt_context* ctx, const float* d_vol, int32_t X, int32_t Y, int32_t Z)
{
    if (!ctx || !d_vol) return fail(VT_E_ARG, "NULL argument");
    if (!ctx->cache) return fail(VT_E_STATE, "no cache");
    if (X <= 0 || Y <= 0 || Z <= 0) return fail(VT_E_ARG, "bad volume dims");
    ctx->cache->host.reset();
    ctx->cache->d_vol = d_vol;
    ctx->cache->vol_dims[0] = X;
    ctx->cache->vol_dims[1] = Y;
    ctx->cache->vol_dims[2] = Z;
    return VT_OK;
}

namespace {

// the pinned host tier of the cache, reading `dataset` of `file` (opened by
// the caller's check)
int attach_host_tier(vt_context* ctx, const char* file, const char* dataset)
{
    BrickCache& c = *ctx->cache;
    std::unique_ptr<HostTier> h(new HostTier);
    h->level_c = c.level_c;
    h->cube = c.cube;
    h->inc = c.inc;
    h->S = c.S;
    h->slot_elems = c.brick_elems;   // the host tier stores bricks back to back
    h->nbricks = c.nbricks;
    h->capacity = std::min<int64_t>(c.cpu_slots, c.nbricks);
    h->file = file;
    h->dataset = dataset;
    size_t bytes = (size_t)h->capacity * c.brick_elems * sizeof(float);
    if (hipHostMalloc((void**)&h->pool, bytes, hipHostMallocMapped) != hipSuccess)
        return fail(VT_E_NOMEM, "pinned host tier allocation failed (" + std::to_string(bytes) + " bytes)");
    HIP_TRY(hipHostGetDevicePointer((void**)&h->pool_dev, h->pool, 0));
    h->slot_of.assign((size_t)c.nbricks, -1);
    h->brick_of.assign((size_t)h->capacity, -1);
    h->ref.assign((size_t)h->capacity, 0);
    h->pin.assign((size_t)h->capacity, 0);
    c.d_vol = nullptr;
    c.host = std::move(h);
    return VT_OK;
}

}  // namespace

int vt_cache_attach_hdf5(vt_context* ctx, const char* file, const char* dataset)
{
    if (!ctx || !file || !dataset) return fail(VT_E_ARG, "NULL argument");
    if (!ctx->cache) return fail(VT_E_STATE, "no cache");
    if (ctx->cache->cpu_slots <= 0) return fail(VT_E_ARG, "an HDF5 source needs a host tier (cpuSlots > 0)");
    {
        H5Volume probe;
        std::string err;
        if (!probe.open(file, dataset, 0, err)) return fail(VT_E_IO, err);
    }
    return attach_host_tier(ctx, file, dataset);
}

int vt_cache_attach_host_from_device(vt_context* ctx, const float* d_vol, int32_t X, int32_t Y, int32_t Z)
{
    if (!ctx || !d_vol) return fail(VT_E_ARG, "NULL argument");
    if (!ctx->cache) return fail(VT_E_STATE, "no cache");
    if (X <= 0 || Y <= 0 || Z <= 0) return fail(VT_E_ARG, "bad volume dims");
    BrickCache& c = *ctx->cache;
    if (!ctx->octree_ready || ctx->max_level < c.level_c || (int)ctx->up_runs.size() <= c.level_c)
        return fail(VT_E_STATE, "upload the octree down to levelC first (the tier holds the bricks with nodes)");
    if ((1 << (ctx->n - c.level_c)) != c.cube) return fail(VT_E_ARG, "cubeDim must equal 2^(nLevels-levelC)");
    if (int rc = set_device(ctx)) return rc;
    const auto& runs = ctx->up_runs[(size_t)c.level_c];
    const uint64_t base = 1ull << (3 * c.level_c);
    std::vector<int64_t> codes;
    for (size_t j = 0; j + 1 < runs.size(); j += 2)
        for (uint64_t code = runs[j]; code <= runs[j + 1]; ++code) codes.push_back((int64_t)code);
    const int64_t K = (int64_t)codes.size();
    if (K > c.cpu_slots) return fail(VT_E_ARG, "the host tier cannot hold every brick with octree nodes");
    std::unique_ptr<HostTier> h(new HostTier);
    h->level_c = c.level_c;
    h->n = ctx->n;
    h->cube = c.cube;
    h->inc = c.inc;
    h->S = c.S;
    h->slot_elems = c.brick_elems;   // bricks back to back
    h->nbricks = c.nbricks;
    h->capacity = std::max<int64_t>(K, 1);
    const size_t bytes = (size_t)h->capacity * c.brick_elems * sizeof(float);
    if (hipHostMalloc((void**)&h->pool, bytes, hipHostMallocMapped) != hipSuccess)
        return fail(VT_E_NOMEM, "pinned host tier allocation failed (" + std::to_string(bytes) + " bytes)");
    HIP_TRY(hipHostGetDevicePointer((void**)&h->pool_dev, h->pool, 0));
    h->slot_of.assign((size_t)c.nbricks, -1);
    h->brick_of.assign((size_t)h->capacity, -1);
    h->ref.assign((size_t)h->capacity, 0);
    h->pin.assign((size_t)h->capacity, 0);
    for (int64_t k = 0; k < K; ++k) {
        h->slot_of[(size_t)(codes[(size_t)k] - (int64_t)base)] = (int32_t)k;
        h->brick_of[(size_t)k] = codes[(size_t)k] - (int64_t)base;
    }
    // assembled on the device (zero-padded halo boxes), a batch at a time,
    // then copied to the pinned slots
    const int64_t B = std::min<int64_t>(std::max<int64_t>(K, 1), 1024);
    DevBuf<float> stage;
    DevBuf<int64_t> d_codes;
    DevBuf<int32_t> d_slots;
    if (stage.ensure((size_t)B * c.brick_elems) != hipSuccess || d_codes.ensure((size_t)B) != hipSuccess ||
        d_slots.ensure((size_t)B) != hipSuccess)
        return fail(VT_E_NOMEM, "staging of the host tier fill");
    std::vector<int32_t> idx((size_t)B);
    for (int64_t q = 0; q < B; ++q) idx[(size_t)q] = (int32_t)q;
    HIP_TRY(hipMemcpy(d_slots.p, idx.data(), (size_t)B * 4, hipMemcpyHostToDevice));
    for (int64_t k0 = 0; k0 < K; k0 += B) {
        const int cnt = (int)std::min<int64_t>(B, K - k0);
        HIP_TRY(hipMemcpyAsync(d_codes.p, codes.data() + k0, (size_t)cnt * 8, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(launch_bricks_from_volume(d_vol, X, Y, Z, d_codes.p, d_slots.p, cnt, c.level_c, ctx->n, c.cube, c.inc,
                                          c.brick_elems, stage.p, ctx->stream));
        HIP_TRY(hipMemcpyAsync(h->pool + (size_t)k0 * c.brick_elems, stage.p, (size_t)cnt * c.brick_elems * 4,
                               hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    stage.release();
    h->used = K;
    h->full_mode = true;
    h->started = true;
    h->slabs_done = (int)h->side_bricks();
    h->filled = K;
    h->complete.store(true, std::memory_order_release);
    c.n = ctx->n;
    c.d_vol = nullptr;
    c.host = std::move(h);
    return VT_OK;
}

int vt_cache_prefetch_host(vt_context* ctx, int wait)
{
    if (!ctx || !ctx->cache || !ctx->cache->host) return fail(VT_E_STATE, "no HDF5-backed cache");
    if (!ctx->octree_ready) return fail(VT_E_STATE, "upload the octree first (brick geometry needs nLevels)");
    HostTier& h = *ctx->cache->host;
    ctx->cache->n = ctx->n;
    if (int rc = host_tier_start(ctx)) return rc;
    if (wait && h.full_mode) {
        std::unique_lock<std::mutex> lk(h.m);
        h.cv.wait(lk, [&] { return h.reader_failed || h.complete.load(std::memory_order_acquire); });
        if (h.reader_failed) return fail(VT_E_IO, "host tier reader: " + h.reader_err);
    }
    return VT_OK;
}

int vt_cache_load_bricks(vt_context* ctx, const uint64_t* codes, int64_t count)
{
    if (!ctx || (count && !codes)) return fail(VT_E_ARG, "NULL argument");
    if (!ctx->cache) return fail(VT_E_STATE, "no cache");
    if (!ctx->octree_ready) return fail(VT_E_STATE, "no octree");
    BrickCache& c = *ctx->cache;
    if (!c.d_vol && !c.host) return fail(VT_E_STATE, "no brick source attached");
    if (int rc = set_device(ctx)) return rc;
    c.n = ctx->n;
    if (int rc = host_tier_start(ctx)) return rc;
    std::vector<uint32_t> req;
    uint64_t base = 1ull << (3 * c.level_c);
    for (int64_t i = 0; i < count; ++i) {
        if (codes[i] < base || codes[i] >= 2 * base) return fail(VT_E_ARG, "brick code outside levelC");
        uint64_t b = codes[i] - base;
        if (c.table[(size_t)b] < 0) req.push_back((uint32_t)b);
    }
    std::sort(req.begin(), req.end());
    req.erase(std::unique(req.begin(), req.end()), req.end());
    return service_requests(ctx, req.data(), (int64_t)req.size(), nullptr);
}

int vt_cache_read_brick(vt_context* ctx, uint64_t code, float* out)
{
    if (!ctx || !out) return fail(VT_E_ARG, "NULL argument");
    if (!ctx->cache) return fail(VT_E_STATE, "no cache");
    BrickCache& c = *ctx->cache;
    const uint64_t base = 1ull << (3 * c.level_c);
    if (code < base || code >= 2 * base) return fail(VT_E_ARG, "brick code outside levelC");
    if (int rc = set_device(ctx)) return rc;
    // loads still on the copy stream land (and are published) first
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (int rc = copy_stream_idle(ctx)) return rc;
    const int32_t slot = c.table[(size_t)(code - base)];
    if (slot < 0) return fail(VT_E_STATE, "brick not resident");
    HIP_TRY(hipMemcpy(out, c.d_pool + (size_t)slot * c.slot_elems, c.brick_elems * sizeof(float),
                      hipMemcpyDeviceToHost));
    return VT_OK;
}

int vt_cache_clear(vt_context* ctx)
{
    if (!ctx || !ctx->cache) return fail(VT_E_STATE, "no cache");
    BrickCache& c = *ctx->cache;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (int rc = copy_stream_idle(ctx)) return rc;
    ctx->early_last = INT64_MAX;
    HIP_TRY(hipMemset(c.d_table, 0xff, (size_t)c.nbricks * 4));
    HIP_TRY(hipMemset(c.d_ref, 0, (size_t)c.gpu_slots));
    std::fill(c.table.begin(), c.table.end(), -1);
    std::fill(c.slot_brick.begin(), c.slot_brick.end(), -1);
    std::fill(c.ref.begin(), c.ref.end(), 0);
    c.used = 0;
    c.hand = 0;
    c.free_list.clear();
    return VT_OK;
}

int vt_cache_get_stats(vt_context* ctx, vt_cache_stats* out)
{
    if (!ctx || !out) return fail(VT_E_ARG, "NULL argument");
    if (!ctx->cache) return fail(VT_E_STATE, "no cache");
    BrickCache& c = *ctx->cache;
    memset(out, 0, sizeof *out);
    out->gpu_slots = c.gpu_slots;
    out->gpu_resident = c.resident();
    out->lookups_missed = c.stat_missed;
    out->uploads = c.stat_uploads;
    out->evictions = c.stat_evictions;
    if (c.host) {
        std::lock_guard<std::mutex> g(c.host->m);
        out->cpu_slots = c.host->capacity;
        out->cpu_resident = c.host->full_mode ? c.host->filled
                                              : c.host->used;
        out->file_reads = c.host->file_reads;
    }
    return VT_OK;
}

int vt_set_frame_options(vt_context* ctx, int collect_stats, int record_trace)
{
    if (!ctx) return fail(VT_E_ARG, "ctx is NULL");
    ctx->collect_stats = collect_stats ? 1 : 0;
    ctx->record_trace = record_trace ? 1 : 0;
    return VT_OK;
}

int vt_set_march_options(vt_context* ctx, int skip_empty, int max_defer, int max_defer_late, int persistent)
{
    if (!ctx) return fail(VT_E_ARG, "ctx is NULL");
    if (max_defer < 0 || max_defer > kMaxDefer || max_defer_late < 0 || max_defer_late > kMaxDefer)
        return fail(VT_E_ARG, "max_defer and max_defer_late must be in [0, 16]");
    ctx->skip_empty = skip_empty ? 1 : 0;
    ctx->max_defer = max_defer;
    ctx->max_defer_late = max_defer_late;
    ctx->persistent = persistent ? 1 : 0;
    return VT_OK;
}

int vt_set_wide_march(vt_context* ctx, int64_t lane_budget, int64_t first_pass_budget)
{
    if (!ctx) return fail(VT_E_ARG, "ctx is NULL");
    if (lane_budget < -6 || first_pass_budget < -6) return fail(VT_E_ARG, "lane budgets must be >= -6");
    ctx->wide_lanes = lane_budget;
    ctx->wide_first = first_pass_budget;
    return VT_OK;
}

int vt_set_first_pass_tiles(vt_context* ctx, int32_t supertile)
{
    if (!ctx) return fail(VT_E_ARG, "ctx is NULL");
    if (supertile < 1 || supertile > 16) return fail(VT_E_ARG, "supertile must be in [1, 16]");
    ctx->supertile = supertile;
    return VT_OK;
}

int vt_set_first_pass_budget(vt_context* ctx, int32_t nodes)
{
    if (!ctx) return fail(VT_E_ARG, "ctx is NULL");
    if (nodes < 0) return fail(VT_E_ARG, "nodes must be >= 0");
    ctx->first_nodes = nodes;
    return VT_OK;
}

int vt_set_early_loads(vt_context* ctx, int64_t reserve)
{
    if (!ctx) return fail(VT_E_ARG, "ctx is NULL");
    if (reserve < 0) return fail(VT_E_ARG, "reserve must be >= 0");
    ctx->early_reserve = reserve;
    ctx->early_last = INT64_MAX;
    return VT_OK;
}

int vt_get_brick_trace(vt_context* ctx, uint8_t* out, int64_t n)
{
    if (!ctx || !out) return fail(VT_E_ARG, "NULL argument");
    if (!ctx->record_trace || ctx->trace_n == 0) return fail(VT_E_STATE, "trace recording is off");
    if (n != ctx->trace_n) return fail(VT_E_ARG, "n must be 8^levelC");
    HIP_TRY(hipMemcpyAsync(out, ctx->trace.p, (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VT_OK;
}

static int render_frame(vt_context* ctx, const vt_camera* cam, const vt_render_params* p, float* d_rgba,
                        vt_hit* d_hits, vt_frame_stats* stats, const RayState* state = nullptr)
{
    double t_start = now_ms();
    if (!ctx || !p || !d_rgba) return fail(VT_E_ARG, "NULL argument");
    if (int rc = check_camera(cam)) return rc;
    if (!ctx->octree_ready) return fail(VT_E_STATE, "no octree uploaded");
    if (!ctx->cache) return fail(VT_E_STATE, "no brick cache");
    BrickCache& c = *ctx->cache;
    if (!c.d_vol && !c.host) return fail(VT_E_STATE, "no brick source attached");
    if (p->levelO < 0 || p->levelO > ctx->max_level) return fail(VT_E_ARG, "levelO beyond the uploaded octree levels");
    if (p->levelC != c.level_c || p->levelC > p->levelO) return fail(VT_E_ARG, "levelC must match the cache and be <= levelO");
    if ((1 << (ctx->n - c.level_c)) != c.cube) return fail(VT_E_ARG, "cubeDim must equal 2^(nLevels-levelC)");
    if (int rc = set_device(ctx)) return rc;
    c.n = ctx->n;
    if (int rc = host_tier_start(ctx)) return rc;

    const uint32_t R = (uint32_t)cam->W * (uint32_t)cam->H;
    RayState own;
    if (!state) {   // the context's buffers, grown to the frame
        if (ctx->ray_node.ensure(R) != hipSuccess || ctx->ray_phase.ensure(R) != hipSuccess ||
            ctx->list[0].ensure(R) != hipSuccess || ctx->list[1].ensure(R) != hipSuccess ||
            (std::max(ctx->max_defer, ctx->max_defer_late) > 0 &&
             ctx->ray_defer.ensure((size_t)R * (size_t)std::max(ctx->max_defer, ctx->max_defer_late)) != hipSuccess))
            return fail(VT_E_NOMEM, "ray state allocation failed");
        own = RayState{ctx->ray_node.p, ctx->ray_defer.p, {ctx->list[0].p, ctx->list[1].p}, ctx->ray_phase.p};
        state = &own;
    }
    HitRecord* hits = reinterpret_cast<HitRecord*>(d_hits);
    if (!hits) {
        if (ctx->hits_scratch.ensure(R) != hipSuccess) return fail(VT_E_NOMEM, "hit buffer");
        hits = ctx->hits_scratch.p;
    }
    if (ctx->record_trace) {
        if (ctx->trace.ensure((size_t)c.nbricks) != hipSuccess) return fail(VT_E_NOMEM, "trace buffer");
        ctx->trace_n = c.nbricks;
        HIP_TRY(hipMemsetAsync(ctx->trace.p, 0, (size_t)c.nbricks, ctx->stream));
    }
    // a frame that ended with no request, no range error and no statistics
    // left the counters and the request bitmap at zero
    if (!ctx->pass_state_clean || ctx->collect_stats)
        HIP_TRY(hipMemsetAsync(ctx->d_ctr, 0, sizeof(FrameCounters), ctx->stream));
    if (!ctx->pass_state_clean) HIP_TRY(hipMemsetAsync(ctx->req_bits.p, 0, ctx->req_bits.n * 4, ctx->stream));
    ctx->pass_state_clean = false;

    RenderArgs a;
    memset(&a, 0, sizeof a);
    a.cam = to_cam(cam);
    a.oct = make_view(ctx, p->levelO);
    a.level_c = c.level_c;
    a.shift = 3 * (p->levelO - p->levelC);
    a.brick_base = 1ull << (3 * c.level_c);
    a.brick_slot = c.d_table;
    a.pool = c.d_pool;
    a.S = c.S;
    a.inc = c.inc;
    a.slot_elems = c.slot_elems;
    a.iso = p->iso;
    a.lx = p->light[0]; a.ly = p->light[1]; a.lz = p->light[2];
    a.color_scale = p->color_scale;
    a.ray_node = state->node;
    a.ray_phase = state->phase;
    a.ray_defer = std::max(ctx->max_defer, ctx->max_defer_late) > 0 ? state->defer : nullptr;
    a.max_defer = ctx->max_defer;
    a.node_budget = ctx->first_nodes;
    a.supertile = ctx->supertile;
    a.req_bits = ctx->req_bits.p;
    a.req_list = ctx->req_list.p;
    a.slot_ref = c.d_ref;
    a.brick_trace = ctx->record_trace ? ctx->trace.p : nullptr;
    a.submm = ctx->skip_empty ? c.d_submm : nullptr;
    a.G = c.G;
    a.sides = ctx->skip_empty ? c.d_sides : nullptr;
    a.side_words = c.side_words;
    a.ctr = ctx->d_ctr;
    a.rgba = reinterpret_cast<float4*>(d_rgba);
    a.hits = hits;
    a.tiles_x = (cam->W + 15) / 16;
    a.tiles_total = a.tiles_x * ((cam->H + 15) / 16);
    a.collect_stats = ctx->collect_stats;

    vt_frame_stats st;
    memset(&st, 0, sizeof st);
    double ms_kernel = 0.0, ms_service = 0.0;
    int cur = 0;
    a.in_list = nullptr;
    a.in_count = 0;
    a.out_list = state->list[cur];
    // bricks still arriving from a prefetch land before the first pass
    if (int rc = copy_stream_idle(ctx)) return rc;
    if (ctx->skip_empty) {
        if (int rc = ensure_sides(ctx, p->iso)) return rc;
    }
    a.side_iso = c.side_iso;
    // free slots for the loads issued while the first pass runs: at most as
    // many as the previous frame's first pass requested, so that a frame
    // whose bricks are all resident evicts nothing
    const bool early = ctx->early_reserve > 0 && !ctx->persistent;
    if (early) {
        if (int rc = make_reserve(ctx, std::min(ctx->early_reserve, ctx->early_last))) return rc;
    }
    uint32_t* rq = ctx->req_list.p;
    a.next = ctx->d_next.p;
    for (int pass = 0;; ++pass) {
        if (ctx->persistent) HIP_TRY(hipMemsetAsync(ctx->d_next.p, 0, 8 * sizeof(uint32_t), ctx->stream));
        HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
        // 2^lg lanes per ray within the pass's lane budget
        int lg = 0;
        const int64_t budget = a.in_list ? ctx->wide_lanes : ctx->wide_first;
        const int64_t rays = a.in_list ? (int64_t)a.in_count : (int64_t)R;
        if (budget > 0)
            while (lg < 6 && (rays << (lg + 1)) <= budget) ++lg;
        else if (budget < 0)
            lg = (int)-budget;
        if (ctx->persistent)
            HIP_TRY(launch_render_persistent(a, ctx->persistent_wgs, ctx->stream));
        else if (lg > 0)
            HIP_TRY(launch_render_wide(a, lg, ctx->stream));
        else
            HIP_TRY(launch_render_pass(a, ctx->stream));
        HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
        // the counters and the reference bytes (for the service) follow the
        // pass on its stream, queued before the host polls
        HIP_TRY(hipMemcpyAsync(ctx->h_ctr, ctx->d_ctr, sizeof(FrameCounters), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(c.h_ref.p, c.d_ref, (size_t)c.gpu_slots, hipMemcpyDeviceToHost, ctx->stream));
        if (early && pass == 0) {
            int64_t loaded = 0, misses = 0;
            if (int rc = load_during_pass(ctx, ctx->ev1, &loaded, &misses)) return rc;
            st.bricks_uploaded += loaded;
            st.bricks_early += loaded;
            st.host_tier_misses += misses;
            // the bricks copied beside the pass become visible to the next one
            if (int rc = publish_loads(ctx, ctx->stream)) return rc;
        }
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ms_kernel += ms;
        st.passes = pass + 1;
        const uint32_t parked = ctx->h_ctr->parked_count;
        const uint32_t nreq = ctx->h_ctr->req_count;
        if (nreq > (uint32_t)c.nbricks) return fail(VT_E_STATE, "request list overflow");
        if (pass == 0) ctx->early_last = nreq;
        st.bricks_requested += nreq;
        // every entry below nreq was stored before the kernel ended
        std::vector<uint32_t>& left = ctx->req_left;
        left.clear();
        for (uint32_t i = 0; i < nreq; ++i) {
            uint32_t v = __atomic_load_n(&rq[i], __ATOMIC_ACQUIRE);
            if (v == 0) return fail(VT_E_STATE, "request entry not visible after the pass");
            left.push_back(v - 1);
        }
        memset(rq, 0, (size_t)nreq * 4);
        if (parked == 0) {
            ctx->pass_state_clean = nreq == 0 && ctx->h_ctr->err == 0;
            break;
        }
        st.rays_parked += parked;
        double t0 = now_ms();
        // the service's last kernel clears the request bitmap, the request
        // and parked counts and the reference bytes for the next pass
        const PassReset reset{ctx->req_bits.p, (int64_t)ctx->req_bits.n, &ctx->d_ctr->req_count, 2, c.d_ref,
                              c.gpu_slots};
        if (!left.empty()) {
            int rc = service_requests(ctx, left.data(), (int64_t)left.size(), &st, &reset);
            if (rc != VT_OK) return rc;
        } else {
            // rays parked on their walk budget only: keep the reference bits
            // of the bricks this pass marched before the service clears them
            if (int rc = fold_ref_bits(ctx, true)) return rc;
            HIP_TRY(launch_finish_service(c.d_table, nullptr, nullptr, 0, reset, ctx->stream));
        }
        ms_service += now_ms() - t0;
        a.in_list = state->list[cur];
        a.in_count = parked;
        a.max_defer = ctx->max_defer_late;   // few rays remain: let them walk further per pass
        a.node_budget = 0;                   // and to the end of their walk
        cur ^= 1;
        a.out_list = state->list[cur];
        if (pass > 1000000) return fail(VT_E_STATE, "frame did not converge");
    }
    c.epoch++;
    c.protect_from = c.epoch;
    st.err = (int32_t)ctx->h_ctr->err;
    st.samples = ctx->h_ctr->samples;
    st.nodes = ctx->h_ctr->nodes;
    st.ms_kernel = ms_kernel;
    st.ms_service = ms_service;
    st.ms_total = now_ms() - t_start;
    if (stats) *stats = st;
    if (st.err) return fail(VT_E_RANGE, "a sample fell outside its brick slot (halo too small for this geometry)");
    return VT_OK;
}

int vt_render_frame(vt_context* ctx, const vt_camera* cam, const vt_render_params* p, float* d_rgba,
                    vt_hit* d_hits, vt_frame_stats* stats)
{
    const int rc = render_frame(ctx, cam, p, d_rgba, d_hits, stats);
    if (rc != VT_OK && ctx) {
        // a frame that fails part-way (a brick source that fails while the
        // first pass runs, say) leaves no kernel or copy running: the caller
        // may free its buffers or the cache at once
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamSynchronize(ctx->cstream);
    }
    return rc;
}

int vt_render_frame_noalloc(vt_context* ctx, const vt_camera* cam, const vt_render_params* p, float* d_rgba,
                            vt_hit* d_hits, vt_frame_stats* stats)
{
    if (!ctx || !p || !d_rgba || !d_hits) return fail(VT_E_ARG, "NULL argument (d_hits is required)");
    if (int rc = check_camera(cam)) return rc;
    const uint64_t R = (uint64_t)cam->W * (uint64_t)cam->H;
    if (R > kScratchRays)
        return fail(VT_E_ARG, "a frame of " + std::to_string(R) + " rays exceeds the allocation-free maximum of " +
                                  std::to_string(kScratchRays));
    if (std::max(ctx->max_defer, ctx->max_defer_late) > kScratchDefer)
        return fail(VT_E_ARG, "max_defer and max_defer_late must be <= 2 for allocation-free frames");
    if (ctx->record_trace) return fail(VT_E_STATE, "the visited-brick trace allocates: turn it off");
    if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch[ctx->device])
        return fail(VT_E_STATE, "no frame scratch on this device");
    uint8_t* base = g_scratch[ctx->device];
    const size_t M = kScratchRays;
    const RayState st{reinterpret_cast<uint64_t*>(base), reinterpret_cast<uint64_t*>(base + 8 * M),
                      {reinterpret_cast<uint32_t*>(base + (8 + 8 * kScratchDefer) * M),
                       reinterpret_cast<uint32_t*>(base + (12 + 8 * kScratchDefer) * M)},
                      base + (16 + 8 * kScratchDefer) * M};
    // calls on one device share the array: one frame at a time (the frame
    // has waited for its last pass when render_frame returns)
    std::lock_guard<std::mutex> g(g_scratch_use[ctx->device]);
    const int rc = render_frame(ctx, cam, p, d_rgba, d_hits, stats, &st);
    if (rc != VT_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamSynchronize(ctx->cstream);
    }
    return rc;
}

int vt_trace_nodes(vt_context* ctx, const vt_camera* cam, int32_t finalLevel, int32_t maxNodes, int32_t capTotal,
                   uint64_t* d_nodes, int32_t* d_counts)
{
    if (!ctx || !d_counts || (maxNodes > 0 && !d_nodes)) return fail(VT_E_ARG, "NULL argument");
    if (int rc = check_camera(cam)) return rc;
    if (!ctx->octree_ready) return fail(VT_E_STATE, "no octree uploaded");
    if (finalLevel < 0 || finalLevel > ctx->max_level) return fail(VT_E_ARG, "finalLevel beyond uploaded levels");
    if (maxNodes < 0 || capTotal < 0) return fail(VT_E_ARG, "negative sizes");
    if (int rc = set_device(ctx)) return rc;
    TraceArgs a;
    a.cam = to_cam(cam);
    a.oct = make_view(ctx, finalLevel);
    a.max_nodes = maxNodes;
    a.cap_total = capTotal;
    a.nodes = d_nodes;
    a.counts = d_counts;
    HIP_TRY(launch_trace_nodes(a, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VT_OK;
}

int vt_file_open(const char* file, const char* dataset, vt_file** out, int64_t dims[3])
{
    if (!file || !dataset || !out) return fail(VT_E_ARG, "NULL argument");
    std::unique_ptr<vt_file> f(new vt_file);
    std::string err;
    if (!f->vol.open(file, dataset, 64ull << 20, err)) return fail(VT_E_IO, err);
    if (dims)
        for (int k = 0; k < 3; ++k) dims[k] = f->vol.dims()[k];
    *out = f.release();
    return VT_OK;
}

int vt_file_close(vt_file* f)
{
    delete f;
    return VT_OK;
}

int vt_file_read_box(vt_file* f, const int32_t s[3], const int32_t e[3], float* out)
{
    if (!f || !s || !e || !out) return fail(VT_E_ARG, "NULL argument");
    int ss[3] = {s[0], s[1], s[2]}, ee[3] = {e[0], e[1], e[2]};
    std::string err;
    if (!f->vol.read_box(ss, ee, out, err)) return fail(VT_E_IO, err);
    return VT_OK;
}

int vt_file_create(const char* file, const char* dataset, const float* data, int64_t X, int64_t Y, int64_t Z,
                   int32_t chunk)
{
    if (!file || !dataset || !data) return fail(VT_E_ARG, "NULL argument");
    if (X <= 0 || Y <= 0 || Z <= 0 || chunk < 0) return fail(VT_E_ARG, "bad dimensions");
    std::string err;
    if (!h5_create_volume(file, dataset, data, X, Y, Z, chunk, err)) return fail(VT_E_IO, err);
    return VT_OK;
}

int vt_file_cut(const char* file, const char* dataset, const char* out_file, int32_t x, int32_t y, int32_t z,
                int32_t dim)
{
    if (!file || !dataset || !out_file) return fail(VT_E_ARG, "NULL argument");
    if (dim <= 0 || dim > 4096) return fail(VT_E_ARG, "dim must be in [1, 4096]");
    // the box must stay in int32 coordinates, as the box read takes them
    const int64_t lim = (int64_t)INT32_MAX - dim;
    if (x > lim || y > lim || z > lim) return fail(VT_E_ARG, "box end overflows");
    vt_file* f = nullptr;
    if (int rc = vt_file_open(file, dataset, &f, nullptr)) return rc;
    std::unique_ptr<vt_file> own(f);
    std::vector<float> box;
    try {
        box.resize((size_t)dim * dim * dim);
    } catch (const std::bad_alloc&) {
        return fail(VT_E_NOMEM, "cut box");
    }
    const int32_t s3[3] = {x, y, z}, e3[3] = {x + dim, y + dim, z + dim};
    if (int rc = vt_file_read_box(f, s3, e3, box.data())) return rc;
    return vt_file_create(out_file, "value", box.data(), dim, dim, dim, 0);
}

}  // extern "C"

// ============================================================ builder ABI

namespace {

// CreateOctree.cu:308-319 (nLevels from the largest extent, float log2)
int tree_levels(int X, int Y, int Z)
{
    int dimension;
    if (X > Y && X > Z) dimension = X;
    else if (Y > Z) dimension = Y;
    else dimension = Z;
    float aux = logf((float)dimension) / logf(2.0f);
    float aux2 = aux - floorf(aux);
    return aux2 > 0.0 ? (int)(aux + 1) : (int)aux;
}

}  // namespace

extern "C" {

int vt_octree_build(vt_context* ctx, const float* d_vol, int32_t X, int32_t Y, int32_t Z, float iso,
                    int32_t literal, int32_t maxLevel, int32_t* nLevels_out, int64_t* sizes_out)
{
    if (!ctx || !d_vol) return fail(VT_E_ARG, "NULL argument");
    if (X <= 0 || Y <= 0 || Z <= 0) return fail(VT_E_ARG, "bad volume dims");
    const int n = tree_levels(X, Y, Z);
    if (n < 2 || n > 12) return fail(VT_E_ARG, "the GPU builder needs 2 <= nLevels <= 12");
    if (maxLevel < 0 || maxLevel > n || maxLevel > 9) return fail(VT_E_ARG, "bad maxLevel");
    if (int rc = set_device(ctx)) return rc;
    hipStream_t s = ctx->stream;
    // bitmaps of levels 2..n, 8^l bits each
    std::vector<uint64_t> woff(n + 1, 0);
    uint64_t total = 0;
    for (int l = 2; l <= n; ++l) { woff[l] = total; total += (1ull << (3 * l)) / 64; }
    DevBuf<uint64_t> bits;
    if (bits.ensure(total) != hipSuccess) return fail(VT_E_NOMEM, "octree bitmap allocation failed");
    HIP_TRY(launch_leaf_bits(d_vol, X, Y, Z, iso, (1ull << (3 * n)) / 64, bits.p + woff[n], s));
    for (int l = n - 1; l >= 2; --l) {
        uint64_t nw = (1ull << (3 * l)) / 64;
        HIP_TRY(launch_parent_bits(bits.p + woff[l + 1], nw, bits.p + woff[l], s));
        if (literal) HIP_TRY(launch_clear_top_bit(bits.p + woff[l] + nw - 1, s));
    }
    // levels 1 and 0 from the single word of level 2
    uint64_t w2 = 0;
    HIP_TRY(hipMemcpyAsync(&w2, bits.p + woff[2], 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    uint32_t lvl1 = 0;
    for (int k = 0; k < 8; ++k)
        if ((w2 >> (8 * k)) & 0xffull) lvl1 |= 1u << k;
    if (literal && n - 1 >= 1) lvl1 &= 0x7fu;   // last node of level 1 (code 15)
    const bool root = lvl1 != 0;                // CreateOctree.cu:532-540
    // run-length encode every level
    std::vector<std::vector<uint64_t>> runs((size_t)n + 1);
    if (root) runs[0] = {1, 1};
    for (int k = 0; k < 8;) {   // level 1 (8 nodes) on the host
        if (!((lvl1 >> k) & 1)) { ++k; continue; }
        int e = k;
        while (e + 1 < 8 && ((lvl1 >> (e + 1)) & 1)) ++e;
        runs[1].push_back(8 + (uint64_t)k);
        runs[1].push_back(8 + (uint64_t)e);
        k = e + 1;
    }
    DevBuf<uint32_t> bsum;
    DevBuf<uint64_t> boff, dout;
    for (int l = 2; l <= n; ++l) {
        uint64_t nbits = 1ull << (3 * l), nw = nbits / 64, nb = run_blocks(nw);
        if (bsum.ensure(nb) != hipSuccess || boff.ensure(nb) != hipSuccess)
            return fail(VT_E_NOMEM, "scan buffers");
        HIP_TRY(launch_run_counts(bits.p + woff[l], nw, nbits, bsum.p, s));
        std::vector<uint32_t> hs(nb);
        std::vector<uint64_t> ho(nb);
        HIP_TRY(hipMemcpyAsync(hs.data(), bsum.p, nb * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        uint64_t acc = 0;
        for (uint64_t b = 0; b < nb; ++b) { ho[b] = acc; acc += hs[b]; }
        if (acc == 0) continue;
        if (dout.ensure(acc) != hipSuccess) return fail(VT_E_NOMEM, "run buffer");
        HIP_TRY(hipMemcpyAsync(boff.p, ho.data(), nb * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(launch_run_write(bits.p + woff[l], nw, nbits, 1ull << (3 * l), boff.p, dout.p, s));
        runs[(size_t)l].resize(acc);
        HIP_TRY(hipMemcpyAsync(runs[(size_t)l].data(), dout.p, acc * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    bits.release();
    bsum.release();
    boff.release();
    dout.release();
    // install levels 0..maxLevel as the traversal octree
    std::vector<uint64_t> flat;
    std::vector<int32_t> sizes((size_t)maxLevel + 1);
    for (int l = 0; l <= maxLevel; ++l) {
        sizes[(size_t)l] = (int32_t)runs[(size_t)l].size();
        flat.insert(flat.end(), runs[(size_t)l].begin(), runs[(size_t)l].end());
    }
    // an isovalue outside the volume's range gives empty run lists
    static const uint64_t no_runs = 0;
    int rc = vt_octree_upload(ctx, flat.empty() ? &no_runs : flat.data(), sizes.data(), n, maxLevel);
    if (rc != VT_OK) return rc;
    ctx->built_runs = std::move(runs);
    ctx->built_iso = iso;
    ctx->built_n = n;
    ctx->built_dim = 1 << n;
    ctx->built_real[0] = X; ctx->built_real[1] = Y; ctx->built_real[2] = Z;
    if (nLevels_out) *nLevels_out = n;
    if (sizes_out)
        for (int l = 0; l <= n; ++l) sizes_out[l] = (int64_t)ctx->built_runs[(size_t)l].size();
    return VT_OK;
}

int vt_octree_get_runs(vt_context* ctx, int32_t level, uint64_t* out, int64_t count)
{
    if (!ctx || (count && !out)) return fail(VT_E_ARG, "NULL argument");
    if (ctx->built_runs.empty()) return fail(VT_E_STATE, "no octree built");
    if (level < 0 || level > ctx->built_n) return fail(VT_E_ARG, "no such level");
    const auto& r = ctx->built_runs[(size_t)level];
    if ((size_t)count != r.size()) return fail(VT_E_ARG, "count must equal the level's run-list size");
    if (count) memcpy(out, r.data(), r.size() * 8);
    return VT_OK;
}

int vt_octree_write_file(vt_context* ctx, const char* path)
{
    if (!ctx || !path) return fail(VT_E_ARG, "NULL argument");
    if (ctx->built_runs.empty()) return fail(VT_E_STATE, "no octree built");
    FILE* f = fopen(path, "wb");
    if (!f) return fail(VT_E_IO, std::string("cannot create ") + path);
    int32_t magic = 919278872, n = ctx->built_n, dim = ctx->built_dim;
    bool ok = fwrite(&magic, 4, 1, f) == 1 && fwrite(&ctx->built_iso, 4, 1, f) == 1 &&
              fwrite(&dim, 4, 1, f) == 1 && fwrite(ctx->built_real, 4, 3, f) == 3 && fwrite(&n, 4, 1, f) == 1;
    for (int l = n; ok && l >= 0; --l) {
        const auto& r = ctx->built_runs[(size_t)l];
        int32_t ne = (int32_t)r.size();
        ok = fwrite(&ne, 4, 1, f) == 1 && (ne == 0 || fwrite(r.data(), 8, r.size(), f) == r.size());
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) return fail(VT_E_IO, std::string("write failed: ") + path);
    return VT_OK;
}

int vt_volume_generate(vt_context* ctx, const vt_gen_params* p, int64_t X, int64_t Y, int64_t Z, int64_t x0,
                       int64_t nx, float* d_out)
{
    if (!ctx || !p || !d_out) return fail(VT_E_ARG, "NULL argument");
    if (X <= 0 || Y <= 0 || Z <= 0 || x0 < 0 || nx <= 0 || x0 + nx > X) return fail(VT_E_ARG, "bad extents");
    if (p->kind < VT_GEN_BLOBS || p->kind > VT_GEN_FOURIER) return fail(VT_E_ARG, "unknown generator");
    if (p->count < 0 || (p->count > 0 && !p->shapes)) return fail(VT_E_ARG, "shapes missing");
    if (p->kind != VT_GEN_BLOBS && p->kind != VT_GEN_FOURIER && (p->octaves < 0 || p->octaves > 16))
        return fail(VT_E_ARG, "octaves must be in [0,16]");
    if (int rc = set_device(ctx)) return rc;
    int stride = p->kind == VT_GEN_SDF ? 8 : 5;
    DevBuf<float> shp;
    if (shp.ensure((size_t)std::max(1, p->count * stride)) != hipSuccess) return fail(VT_E_NOMEM, "shapes");
    if (p->count)
        HIP_TRY(hipMemcpyAsync(shp.p, p->shapes, (size_t)p->count * stride * 4, hipMemcpyHostToDevice, ctx->stream));
    GenParams g;
    g.kind = p->kind;
    g.seed = p->seed;
    g.X = X; g.Y = Y; g.Z = Z;
    g.n = p->n > 0 ? p->n : (int)X;
    g.count = p->count;
    g.shapes = shp.p;
    g.octaves = p->octaves;
    g.base_freq = p->base_freq;
    g.gain = p->gain;
    g.lacunarity = p->lacunarity;
    g.scale = p->scale;
    HIP_TRY(launch_generate(d_out, g, x0, nx, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    shp.release();
    return VT_OK;
}

int vt_volume_write_bricks_hdf5(vt_context* ctx, const float* d_vol, int64_t X, int64_t Y, int64_t Z, int32_t levelC,
                                int32_t cubeDim, int32_t cubeInc, const uint64_t* codes, int64_t ncodes,
                                const char* file, const char* dataset)
{
    if (!ctx || !d_vol || !codes || !file || !dataset) return fail(VT_E_ARG, "NULL argument");
    if (X <= 0 || Y <= 0 || Z <= 0 || X > INT32_MAX || Y > INT32_MAX || Z > INT32_MAX)
        return fail(VT_E_ARG, "bad dimensions");
    if (levelC < 0 || levelC > 9 || cubeDim <= 0 || (cubeDim & (cubeDim - 1)) || cubeInc < 0 || ncodes <= 0)
        return fail(VT_E_ARG, "bad brick layout");
    const uint64_t lo = 1ull << (3 * levelC), hi = 1ull << (3 * levelC + 3);
    for (int64_t k = 0; k < ncodes; ++k)
        if (codes[k] < lo || codes[k] >= hi || (k > 0 && codes[k] <= codes[k - 1]))
            return fail(VT_E_ARG, "codes must be ascending codes of level levelC");
    int n = levelC;
    while ((1 << (n - levelC)) < cubeDim) ++n;
    if (int rc = set_device(ctx)) return rc;
    const int S = cubeDim + 2 * cubeInc;
    const size_t elems = (size_t)S * S * S;
    const int64_t dims[3] = {X, Y, Z};
    std::string err;
    H5BrickWriter w;
    if (!w.create(file, dataset, codes, ncodes, levelC, cubeDim, cubeInc, dims, err)) return fail(VT_E_IO, err);
    // batches of bricks gathered on the device (k_bricks_from_volume) and
    // written as consecutive records
    const int64_t per = std::max<int64_t>(1, std::min<int64_t>(ncodes, (int64_t)((256ull << 20) / (elems * 4))));
    DevBuf<float> d_bricks;
    DevBuf<int64_t> d_codes;
    DevBuf<int32_t> d_slots;
    PinBuf<float> h_bricks;
    if (d_bricks.ensure((size_t)per * elems) != hipSuccess || d_codes.ensure((size_t)per) != hipSuccess ||
        d_slots.ensure((size_t)per) != hipSuccess || h_bricks.ensure((size_t)per * elems) != hipSuccess)
        return fail(VT_E_NOMEM, "brick staging");
    std::vector<int32_t> slots((size_t)per);
    for (int64_t i = 0; i < per; ++i) slots[(size_t)i] = (int32_t)i;
    HIP_TRY(hipMemcpyAsync(d_slots.p, slots.data(), (size_t)per * 4, hipMemcpyHostToDevice, ctx->stream));
    for (int64_t k0 = 0; k0 < ncodes; k0 += per) {
        const int64_t nb = std::min(per, ncodes - k0);
        HIP_TRY(hipMemcpyAsync(d_codes.p, codes + k0, (size_t)nb * 8, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(launch_bricks_from_volume(d_vol, (int)X, (int)Y, (int)Z, d_codes.p, d_slots.p, (int)nb, levelC, n,
                                          cubeDim, cubeInc, elems, d_bricks.p, ctx->stream));
        HIP_TRY(hipMemcpyAsync(h_bricks.p, d_bricks.p, (size_t)nb * elems * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (!w.write(k0, nb, h_bricks.p, err)) return fail(VT_E_IO, err);
    }
    if (!w.close(err)) return fail(VT_E_IO, err);
    return VT_OK;
}

int vt_cache_attach_hdf5_bricks(vt_context* ctx, const char* file, const char* dataset)
{
    if (!ctx || !file || !dataset) return fail(VT_E_ARG, "NULL argument");
    if (!ctx->cache) return fail(VT_E_STATE, "no cache");
    BrickCache& c = *ctx->cache;
    std::string err;
    H5BrickReader rd;
    if (!rd.open(file, dataset, err)) return fail(VT_E_IO, err);
    if (rd.level_c != c.level_c || rd.cube != c.cube || rd.inc != c.inc)
        return fail(VT_E_ARG, "the brick file's levelC/cubeDim/cubeInc (" + std::to_string(rd.level_c) + "/" +
                                  std::to_string(rd.cube) + "/" + std::to_string(rd.inc) + ") differ from the cache's");
    if (c.cpu_slots <= 0) return fail(VT_E_ARG, "an HDF5 source needs a host tier (cpuSlots > 0)");
    if (int rc = attach_host_tier(ctx, file, dataset)) return rc;
    HostTier& h = *c.host;
    h.bricked = true;
    h.rec_of.assign((size_t)c.nbricks, -1);
    const uint64_t base = 1ull << (3 * c.level_c);
    for (size_t k = 0; k < rd.codes.size(); ++k) h.rec_of[(size_t)(rd.codes[k] - base)] = (int64_t)k;
    return VT_OK;
}

int vt_volume_write_hdf5(vt_context* ctx, const float* d_vol, int64_t X, int64_t Y, int64_t Z, const char* file,
                         const char* dataset, int32_t chunk)
{
    if (!ctx || !d_vol || !file || !dataset) return fail(VT_E_ARG, "NULL argument");
    if (X <= 0 || Y <= 0 || Z <= 0 || chunk < 0) return fail(VT_E_ARG, "bad dimensions");
    if (int rc = set_device(ctx)) return rc;
    std::string err;
    H5Writer w;
    if (!w.create(file, dataset, X, Y, Z, chunk, err)) return fail(VT_E_IO, err);
    // slabs of whole chunk rows, ~256 MB each, double-buffered in pinned memory
    int64_t step = chunk > 0 ? chunk : 1;
    int64_t per = std::max<int64_t>(step, ((256ll << 20) / (Y * Z * 4)) / step * step);
    PinBuf<float> hb[2];
    if (hb[0].ensure((size_t)(per * Y * Z)) != hipSuccess || hb[1].ensure((size_t)(per * Y * Z)) != hipSuccess)
        return fail(VT_E_NOMEM, "pinned slab buffers");
    int cur = 0;
    int64_t x0 = 0, nx0 = std::min(per, X);
    HIP_TRY(hipMemcpyAsync(hb[cur].p, d_vol, (size_t)(nx0 * Y * Z) * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    while (x0 < X) {
        int64_t nx = std::min(per, X - x0);
        int64_t x1 = x0 + nx;
        if (x1 < X) {   // next slab's copy overlaps this slab's file write
            int64_t nn = std::min(per, X - x1);
            HIP_TRY(hipMemcpyAsync(hb[cur ^ 1].p, d_vol + (size_t)x1 * Y * Z, (size_t)(nn * Y * Z) * 4,
                                   hipMemcpyDeviceToHost, ctx->stream));
        }
        if (!w.write_slab(x0, nx, hb[cur].p, err)) return fail(VT_E_IO, err);
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        cur ^= 1;
        x0 = x1;
    }
    if (!w.close(err)) return fail(VT_E_IO, err);
    return VT_OK;
}

}  // extern "C"

// ======================================================= staged protocol

extern "C" {

int vt_stage_reset(vt_context* ctx, int64_t numRays)
{
    if (!ctx || numRays <= 0) return fail(VT_E_ARG, "bad arguments");
    if (int rc = set_device(ctx)) return rc;
    if (ctx->stage_trav.ensure((size_t)numRays) != hipSuccess) return fail(VT_E_NOMEM, "traversal state");
    HIP_TRY(hipMemsetAsync(ctx->stage_trav.p, 0, (size_t)numRays, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->stage_rays = numRays;
    return VT_OK;
}

int vt_stage_traverse(vt_context* ctx, const vt_camera* cam, int32_t finalLevel, vt_visible_cube* d_cubes)
{
    if (!ctx || !d_cubes) return fail(VT_E_ARG, "NULL argument");
    if (int rc = check_camera(cam)) return rc;
    if (!ctx->octree_ready) return fail(VT_E_STATE, "no octree uploaded");
    if (finalLevel < 0 || finalLevel > ctx->max_level) return fail(VT_E_ARG, "finalLevel beyond uploaded levels");
    int64_t R = (int64_t)cam->W * cam->H;
    if (ctx->stage_rays < R) {
        if (int rc = vt_stage_reset(ctx, R)) return rc;
    }
    if (int rc = set_device(ctx)) return rc;
    StageTraverseArgs a;
    a.cam = to_cam(cam);
    a.oct = make_view(ctx, finalLevel);
    a.cubes = reinterpret_cast<VisibleCube*>(d_cubes);
    a.trav = ctx->stage_trav.p;
    HIP_TRY(launch_stage_traverse(a, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VT_OK;
}

int vt_stage_cache(vt_context* ctx, vt_visible_cube* h_cubes, int64_t num, int32_t levelO, int32_t nLevels)
{
    if (!ctx || (num && !h_cubes)) return fail(VT_E_ARG, "NULL argument");
    if (!ctx->cache) return fail(VT_E_STATE, "no cache");
    BrickCache& c = *ctx->cache;
    if (!c.d_vol && !c.host) return fail(VT_E_STATE, "no brick source attached");
    if (levelO < c.level_c || nLevels < levelO || nLevels > 20) return fail(VT_E_ARG, "bad levels");
    if ((1 << (nLevels - c.level_c)) != c.cube) return fail(VT_E_ARG, "cubeDim must equal 2^(nLevels-levelC)");
    if (int rc = set_device(ctx)) return rc;
    c.n = nLevels;
    if (int rc = host_tier_start(ctx)) return rc;
    const int shift = 3 * (levelO - c.level_c);
    const uint64_t base = 1ull << (3 * c.level_c);
    std::vector<uint32_t> need;
    for (int64_t i = 0; i < num; ++i) {
        uint8_t st = h_cubes[i].state;
        if (st != VT_CUBE && st != VT_NOCACHED) continue;
        uint64_t b = h_cubes[i].id >> shift;
        if (b < base || b >= 2 * base) return fail(VT_E_ARG, "node code outside the octree");
        need.push_back((uint32_t)(b - base));
    }
    std::sort(need.begin(), need.end());
    need.erase(std::unique(need.begin(), need.end()), need.end());
    if (int rc = fold_ref_bits(ctx)) return rc;
    // pin every brick this pass already holds, so that loading the missing
    // ones can never evict a brick another ray of the pass points at
    // (lruCache.cu:254-266 could)
    c.epoch++;
    int64_t pinned = 0;
    std::vector<uint32_t> missing;
    for (uint32_t b : need) {
        int32_t s = c.table[b];
        if (s >= 0) {
            c.pin[(size_t)s] = c.epoch;
            c.ref[(size_t)s] = 1;
            pinned++;
        } else {
            missing.push_back(b);
        }
    }
    int64_t avail = c.gpu_slots - pinned;
    int64_t n_load = std::min<int64_t>((int64_t)missing.size(), avail);
    if (c.host && !c.host->full_mode) n_load = std::min<int64_t>(n_load, c.host->capacity);
    int64_t misses = 0;
    if (int rc = load_bricks(ctx, missing.data(), n_load, &misses, nullptr)) return rc;
    c.stat_missed += (int64_t)missing.size();
    for (int64_t i = 0; i < num; ++i) {
        vt_visible_cube& v = h_cubes[i];
        if (v.state != VT_CUBE && v.state != VT_NOCACHED) continue;
        uint64_t brick = v.id >> shift;
        int32_t s = c.table[(size_t)(brick - base)];
        if (s >= 0 && c.pin[(size_t)s] == c.epoch) {
            v.data = c.d_pool + (size_t)s * c.slot_elems;
            v.state = VT_CACHED;
            v.cubeID = brick;
        } else {
            v.state = VT_NOCACHED;   // retried next pass (lruCache.cu:184-191)
            v.cubeID = 0;
        }
    }
    return VT_OK;
}

int vt_stage_raycast(vt_context* ctx, const vt_camera* cam, const vt_render_params* p, int32_t nLevels,
                     int32_t cubeDim, int32_t cubeInc, vt_visible_cube* d_cubes, float* d_rgba)
{
    if (!ctx || !p || !d_cubes || !d_rgba) return fail(VT_E_ARG, "NULL argument");
    if (int rc = check_camera(cam)) return rc;
    if (nLevels < 1 || nLevels > 20 || p->levelO > nLevels || p->levelC > p->levelO || cubeDim <= 0 || cubeInc < 0)
        return fail(VT_E_ARG, "bad levels or brick geometry");
    if ((1 << (nLevels - p->levelC)) != cubeDim) return fail(VT_E_ARG, "cubeDim must equal 2^(nLevels-levelC)");
    if (int rc = set_device(ctx)) return rc;
    StageRaycastArgs a;
    a.cam = to_cam(cam);
    a.n = nLevels;
    a.level_o = p->levelO;
    a.level_c = p->levelC;
    a.inc = cubeInc;
    a.S = cubeDim + 2 * cubeInc;
    a.iso = p->iso;
    a.lx = p->light[0]; a.ly = p->light[1]; a.lz = p->light[2];
    a.color_scale = p->color_scale;
    a.cubes = reinterpret_cast<VisibleCube*>(d_cubes);
    a.rgba = reinterpret_cast<float4*>(d_rgba);
    a.err = &ctx->d_ctr->err;
    ctx->pass_state_clean = false;
    HIP_TRY(hipMemsetAsync(&ctx->d_ctr->err, 0, 4, ctx->stream));
    HIP_TRY(launch_stage_raycast(a, ctx->stream));
    uint32_t err = 0;
    HIP_TRY(hipMemcpyAsync(&ctx->h_ctr->err, &ctx->d_ctr->err, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    err = ctx->h_ctr->err;
    if (err) return fail(VT_E_RANGE, "a sample fell outside its brick slot (halo too small for this geometry)");
    return VT_OK;
}

}  // extern "C"

extern "C" int vt_render_cube(vt_context* ctx, const vt_camera* cam, float iso, const float light[3],
                              const float* d_cube, const int32_t minBox[3], const int32_t cubeDim[3],
                              int32_t cubeInc, float* d_rgba)
{
    if (!ctx || !light || !d_cube || !minBox || !cubeDim || !d_rgba) return fail(VT_E_ARG, "NULL argument");
    if (int rc = check_camera(cam)) return rc;
    // posToIndex (rayCaster.cu:9) strides every axis by dim.x: only cubic bricks are laid out consistently
    if (cubeDim[0] <= 0 || cubeDim[0] != cubeDim[1] || cubeDim[0] != cubeDim[2] || cubeInc < 0)
        return fail(VT_E_ARG, "renderCube needs a cubic brick");
    if (int rc = set_device(ctx)) return rc;
    CubeArgs a;
    a.cam = to_cam(cam);
    a.iso = iso;
    a.lx = light[0]; a.ly = light[1]; a.lz = light[2];
    a.cube = d_cube;
    a.mnx = minBox[0]; a.mny = minBox[1]; a.mnz = minBox[2];
    a.dim = cubeDim[0];
    a.inc = cubeInc;
    a.rgba = reinterpret_cast<float4*>(d_rgba);
    a.err = &ctx->d_ctr->err;
    ctx->pass_state_clean = false;
    HIP_TRY(hipMemsetAsync(&ctx->d_ctr->err, 0, 4, ctx->stream));
    HIP_TRY(launch_render_cube(a, ctx->stream));
    HIP_TRY(hipMemcpyAsync(&ctx->h_ctr->err, &ctx->d_ctr->err, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->h_ctr->err) return fail(VT_E_RANGE, "a sample fell outside the brick (halo too small)");
    return VT_OK;
}

extern "C" int vt_image_write_ppm(const float* h_rgba, int32_t W, int32_t H, const char* path)
{
    if (!h_rgba || !path) return fail(VT_E_ARG, "NULL argument");
    if (W <= 0 || H <= 0) return fail(VT_E_ARG, "image size must be positive");
    auto byte = [](float v) -> unsigned char {
        float s = v * 255.0f;
        if (!(s > 0.0f)) return 0;   // also NaN
        if (s >= 255.0f) return 255;
        return (unsigned char)s;     // truncation, as the reference's BYTE conversion
    };
    const int width = H, height = W;   // FreeImage_Allocate(H, W, 24)
    std::vector<unsigned char> img((size_t)width * height * 3);
    for (int r = 0; r < height; ++r) {   // PPM rows run top-down, FreeImage's bottom-up
        const int y = height - 1 - r;
        for (int x = 0; x < width; ++x) {
            const float* p = h_rgba + ((size_t)x * W + y) * 4;   // ray i*W + j with i = x, j = y
            unsigned char* q = img.data() + ((size_t)r * width + x) * 3;
            q[0] = byte(p[0]);
            q[1] = byte(p[1]);
            q[2] = byte(p[2]);
        }
    }
    FILE* f = fopen(path, "wb");
    if (!f) return fail(VT_E_IO, std::string("cannot create ") + path);
    fprintf(f, "P6\n%d %d\n255\n", width, height);
    size_t wrote = fwrite(img.data(), 1, img.size(), f);
    int closed = fclose(f);
    if (wrote != img.size() || closed != 0) return fail(VT_E_IO, std::string("writing ") + path);
    return VT_OK;
}

extern "C" int vt_copy_probe(vt_context* ctx, const void* d_src, void* d_dst, size_t bytes)
{
    if (!ctx || !d_src || !d_dst || bytes % 16) return fail(VT_E_ARG, "bad arguments");
    if (int rc = set_device(ctx)) return rc;
    HIP_TRY(launch_copy_probe((const float*)d_src, (float*)d_dst, bytes, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VT_OK;
}
