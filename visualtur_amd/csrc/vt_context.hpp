// vt_context.hpp -- private state of a vt_context (host runtime).
#pragma once
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

#include "visualtur_amd.h"
#include "vt_hdf5.hpp"
#include "vt_internal.hpp"

using namespace vt;

static_assert(sizeof(vt_camera) == sizeof(vt::Camera), "camera layout");
static_assert(sizeof(vt_hit) == sizeof(vt::HitRecord), "hit layout");

namespace vt_rt {
int fail(int code, const std::string& msg);
extern thread_local std::string g_err;   // message of the calling thread's last failure

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(VT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t count)
    {
        if (count <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

template <class T>
struct PinBuf {
    T* p = nullptr;
    size_t n = 0;
    T* dev = nullptr;   // the same memory as seen by kernels
    hipError_t ensure(size_t count, unsigned flags = hipHostMallocDefault)
    {
        if (count <= n && p) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        dev = nullptr;
        n = 0;
        hipError_t e = hipHostMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T), flags | hipHostMallocMapped);
        if (e != hipSuccess) return e;
        e = hipHostGetDevicePointer((void**)&dev, p, 0);
        if (e == hipSuccess) n = count;
        return e;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        dev = nullptr;
        n = 0;
    }
};

// ------------------------------------------------------------ host tier
// Pinned, device-visible brick slots.  When it can hold every brick of the
// cube, or every brick that holds octree nodes (the only ones a ray can
// request), a reader thread fills it slab by slab (x-slabs of bricks, one
// HDF5 hyperslab read per slab); otherwise bricks are read on demand into
// clock-evicted slots.
struct HostTier {
    int level_c = 0, n = 0, cube = 0, inc = 0, S = 0;
    int64_t nbricks = 0, capacity = 0;
    size_t slot_elems = 0;
    float* pool = nullptr;        // pinned host memory
    float* pool_dev = nullptr;    // the same memory as seen by kernels
    std::string file, dataset;
    H5Volume sync_reader;         // on-demand reads (service thread)
    // brick-major source (H5BrickReader layout): record of each brick index,
    // -1 when the file holds none
    bool bricked = false;
    std::vector<int64_t> rec_of;
    H5BrickReader brick_sync;
    int64_t recs_done = 0;        // records the reader has passed (full mode)
    std::vector<int32_t> slot_of; // brick index -> host slot
    std::vector<int64_t> brick_of;
    std::vector<uint8_t> ref;
    std::vector<uint64_t> pin;
    int64_t hand = 0, used = 0;
    // prefetch state
    bool full_mode = false;
    bool started = false;
    std::thread reader;
    std::mutex m;
    std::condition_variable cv;
    int slabs_done = 0;           // x-slabs of bricks fully loaded
    int64_t filled = 0;           // bricks the reader has placed
    std::atomic<bool> complete{false};   // every slab loaded: lookups need no lock
    bool reader_failed = false;
    std::string reader_err;
    int64_t file_reads = 0;

    ~HostTier()
    {
        if (reader.joinable()) reader.join();
        if (pool) (void)hipHostFree(pool);
    }

    int64_t side_bricks() const { return 1ll << level_c; }

    // linear x-major brick index of (bx,by,bz) brick coordinates
    static uint64_t brick_code(int bx, int by, int bz, int level_c)
    {
        return (1ull << (3 * level_c)) | (dilate3((uint32_t)bx) << 2) | (dilate3((uint32_t)by) << 1) |
               dilate3((uint32_t)bz);
    }

    void reader_failure(const std::string& e)
    {
        std::lock_guard<std::mutex> g(m);
        reader_failed = true;
        reader_err = e;
        cv.notify_all();
    }

    // full mode from a brick-major file: records in file order, each read
    // straight into its host slot
    void run_prefetch_bricked()
    {
        H5BrickReader rd;
        std::string err;
        if (!rd.open(file, dataset, err)) return reader_failure(err);
        const uint64_t base = 1ull << (3 * level_c);
        const int64_t K = (int64_t)rd.codes.size();
        int64_t placed = 0;
        for (int64_t k = 0; k < K; ++k) {
            const int64_t bidx = (int64_t)(rd.codes[(size_t)k] - base);
            if (slot_of[(size_t)bidx] >= 0) {
                if (!rd.read(k, pool + (size_t)slot_of[(size_t)bidx] * slot_elems, err)) return reader_failure(err);
                placed++;
            }
            if ((k & 63) == 63 || k + 1 == K) {
                {
                    std::lock_guard<std::mutex> g(m);
                    recs_done = k + 1;
                    file_reads += placed;
                    filled += placed;
                    placed = 0;
                    if (k + 1 == K) complete.store(true, std::memory_order_release);
                }
                cv.notify_all();
            }
        }
    }

    void run_prefetch()
    {
        if (bricked) return run_prefetch_bricked();
        H5Volume vol;
        std::string err;
        if (!vol.open(file, dataset, 0, err)) {
            std::lock_guard<std::mutex> g(m);
            reader_failed = true;
            reader_err = err;
            cv.notify_all();
            return;
        }
        const int64_t B = side_bricks();
        const int64_t ext = B * cube;   // cube extent in voxels
        const int64_t SY = ext + 2 * inc, SZ = ext + 2 * inc;
        std::vector<float> slab((size_t)S * SY * SZ);
        for (int64_t bx = 0; bx < B; ++bx) {
            int s[3] = {(int)(bx * cube - inc), -inc, -inc};
            int e[3] = {(int)(bx * cube - inc + S), (int)(ext + inc), (int)(ext + inc)};
            if (!vol.read_box(s, e, slab.data(), err)) {
                std::lock_guard<std::mutex> g(m);
                reader_failed = true;
                reader_err = err;
                cv.notify_all();
                return;
            }
            int64_t placed = 0;
            for (int64_t by = 0; by < B; ++by)
                for (int64_t bz = 0; bz < B; ++bz) {
                    uint64_t code = brick_code((int)bx, (int)by, (int)bz, level_c);
                    int64_t bidx = (int64_t)(code - (1ull << (3 * level_c)));
                    if (slot_of[bidx] < 0) continue;   // holds no octree node
                    placed++;
                    float* dst = pool + (size_t)slot_of[bidx] * slot_elems;
                    for (int x = 0; x < S; ++x)
                        for (int y = 0; y < S; ++y) {
                            const float* src = slab.data() + ((size_t)x * SY + (size_t)(by * cube + y)) * SZ +
                                               (size_t)(bz * cube);
                            memcpy(dst + ((size_t)x * S + y) * S, src, (size_t)S * sizeof(float));
                        }
                }
            {
                std::lock_guard<std::mutex> g(m);
                slabs_done = (int)bx + 1;
                file_reads += placed;
                filled += placed;
                if (bx + 1 == B) complete.store(true, std::memory_order_release);
            }
            cv.notify_all();
        }
    }

    // runs_c: the octree's level-levelC run list, or null when it is not
    // uploaded (then only a tier that holds every brick is filled ahead)
    int start(int n_levels, const std::vector<uint64_t>* runs_c)
    {
        started = true;
        n = n_levels;
        if (bricked && runs_c) {
            // a brick-major file is read ahead only for the bricks with
            // nodes, and only when it holds a record for each of them
            const uint64_t base = 1ull << (3 * level_c);
            int64_t count = 0;
            for (size_t j = 0; j + 1 < runs_c->size(); j += 2)
                for (uint64_t code = (*runs_c)[j]; code <= (*runs_c)[j + 1]; ++code, ++count)
                    if (rec_of[(size_t)(code - base)] < 0)
                        return fail(VT_E_ARG, "the brick file holds no record for brick code " + std::to_string(code) +
                                                  ", which the octree needs");
            if (count > capacity) return VT_OK;   // on demand
        } else if (bricked) {
            return VT_OK;   // on demand
        }
        if (capacity >= nbricks && !bricked) {
            full_mode = true;
            for (int64_t b = 0; b < nbricks; ++b) { slot_of[b] = (int32_t)b; brick_of[b] = b; }
            used = nbricks;
        } else if (runs_c) {
            const uint64_t base = 1ull << (3 * level_c);
            int64_t count = 0;
            for (size_t j = 0; j + 1 < runs_c->size(); j += 2) count += (int64_t)((*runs_c)[j + 1] - (*runs_c)[j] + 1);
            if (count > capacity) return VT_OK;   // on demand
            full_mode = true;
            int64_t k = 0;
            for (size_t j = 0; j + 1 < runs_c->size(); j += 2)
                for (uint64_t code = (*runs_c)[j]; code <= (*runs_c)[j + 1]; ++code, ++k) {
                    slot_of[(size_t)(code - base)] = (int32_t)k;
                    brick_of[(size_t)k] = (int64_t)(code - base);
                }
            used = count;
        }
        if (full_mode) reader = std::thread([this] { run_prefetch(); });
        return VT_OK;
    }

    // device-visible pointer of brick `bidx`, reading it if needed.
    int get(int64_t bidx, uint64_t epoch, const float** out, int64_t* misses)
    {
        if (full_mode && complete.load(std::memory_order_acquire)) {
            if (slot_of[bidx] < 0) return fail(VT_E_ARG, "requested brick holds no octree node");
            *out = pool_dev + (size_t)slot_of[bidx] * slot_elems;
            return VT_OK;
        }
        if (full_mode && bricked) {
            if (slot_of[bidx] < 0) return fail(VT_E_ARG, "requested brick holds no octree node");
            const int64_t rec = rec_of[(size_t)bidx];
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return reader_failed || recs_done > rec; });
            if (reader_failed) return fail(VT_E_IO, "host tier reader: " + reader_err);
            *out = pool_dev + (size_t)slot_of[bidx] * slot_elems;
            return VT_OK;
        }
        if (full_mode) {
            int32_t code_bx;
            {
                int x, y, z;
                node_min(bidx + (1ull << (3 * level_c)), level_c, level_c, x, y, z);
                code_bx = x;
            }
            if (slot_of[bidx] < 0) return fail(VT_E_ARG, "requested brick holds no octree node");
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return reader_failed || slabs_done > code_bx; });
            if (reader_failed) return fail(VT_E_IO, "host tier reader: " + reader_err);
            *out = pool_dev + (size_t)slot_of[bidx] * slot_elems;
            return VT_OK;
        }
        int32_t s = slot_of[bidx];
        if (s >= 0) {
            ref[s] = 1;
            pin[s] = epoch;
            *out = pool_dev + (size_t)s * slot_elems;
            return VT_OK;
        }
        // the file first: a source that cannot be opened takes no slot
        if (bricked) {
            if (rec_of[(size_t)bidx] < 0)
                return fail(VT_E_ARG, "the brick file holds no record for brick index " + std::to_string(bidx));
            if (!brick_sync.is_open()) {
                std::string err;
                if (!brick_sync.open(file, dataset, err)) return fail(VT_E_IO, err);
            }
        } else if (!sync_reader.is_open()) {
            std::string err;
            if (!sync_reader.open(file, dataset, 64ull << 20, err)) return fail(VT_E_IO, err);
        }
        // choose a host slot: unused first, then clock skipping pinned (a
        // slot emptied by a failed read holds no brick and is taken as is)
        int64_t slot = -1;
        if (used < capacity) {
            slot = used++;
        } else {
            for (int64_t it = 0; it < 3 * capacity; ++it) {
                int64_t c = hand;
                hand = (hand + 1) % capacity;
                if (pin[c] == epoch) continue;
                if (ref[c]) { ref[c] = 0; continue; }
                slot = c;
                break;
            }
            if (slot < 0) return fail(VT_E_STATE, "host tier: every slot is pinned");
        }
        if (brick_of[slot] >= 0) {
            slot_of[(size_t)brick_of[slot]] = -1;
            brick_of[slot] = -1;
        }
        std::string err;
        if (bricked) {
            if (!brick_sync.read(rec_of[(size_t)bidx], pool + (size_t)slot * slot_elems, err)) return fail(VT_E_IO, err);
        } else {
            int x, y, z;
            node_min(bidx + (1ull << (3 * level_c)), level_c, n, x, y, z);
            int s3[3] = {x - inc, y - inc, z - inc};
            int e3[3] = {x - inc + S, y - inc + S, z - inc + S};
            if (!sync_reader.read_box(s3, e3, pool + (size_t)slot * slot_elems, err)) return fail(VT_E_IO, err);
        }
        brick_of[slot] = bidx;
        slot_of[bidx] = (int32_t)slot;
        ref[slot] = 1;
        pin[slot] = epoch;
        file_reads++;
        (*misses)++;
        *out = pool_dev + (size_t)slot * slot_elems;
        return VT_OK;
    }
};

// Pinned, device-visible staging of the loads issued on the copy stream (the
// copy, sub-block and table-scatter kernels read it in place).  Entries are
// appended while the copy stream works and reset only when it is idle.
struct AsyncStage {
    PinBuf<const float*> src;
    PinBuf<int32_t> slots;
    PinBuf<int64_t> codes;
    PinBuf<int64_t> tidx;
    PinBuf<int32_t> tval;
    int64_t used = 0, cap = 0;
    int64_t pub = 0;   // entries below pub have their sub-block min/max and page-table entries written
    hipError_t ensure(int64_t count)
    {
        hipError_t e;
        if ((e = src.ensure((size_t)count)) || (e = slots.ensure((size_t)count)) ||
            (e = codes.ensure((size_t)count)) || (e = tidx.ensure((size_t)count)) || (e = tval.ensure((size_t)count)))
            return e;
        cap = count;
        used = pub = 0;
        return hipSuccess;
    }
    void release()
    {
        src.release(); slots.release(); codes.release(); tidx.release(); tval.release();
        used = cap = pub = 0;
    }
};

// ----------------------------------------------------------- brick cache
// Byte offsets of the staging arrays of a load of n bricks (n table entries
// for the loads and up to n for the evictions): source pointers, brick codes,
// table indices, slots, table values.
struct StageLayout {
    size_t src, codes, tidx, slots, tval, bytes;
    explicit StageLayout(size_t n)
        : src(0), codes(8 * n), tidx(16 * n), slots(32 * n), tval(36 * n), bytes(44 * n) {}
};

struct BrickCache {
    int level_c = 0, cube = 0, inc = 0, S = 0;
    int n = 0;   // octree depth the brick codes refer to (set on first use)
    size_t brick_elems = 0;   // S^3 floats of a brick
    // floats between device slots: S^3 rounded up to 256 bytes, so no cache
    // line holds two slots (a copy into a free slot while a pass runs must
    // not share a line that the pass reads from its neighbour)
    size_t slot_elems = 0;
    int64_t nbricks = 0, gpu_slots = 0, cpu_slots = 0;
    float* d_pool = nullptr;
    int32_t* d_table = nullptr;
    uint8_t* d_ref = nullptr;
    float2* d_submm = nullptr;   // G^3 sub-block (min, max) per slot, null when bricks are not 4k wide
    int G = 0;
    // per slot, the sides the sub-blocks' (min, max) prove for side_iso, 2
    // bits a sub-block, side_words words a slot (a multiple of 32: one
    // 128-byte line per 32^3 brick); NaN side_iso: not computed yet
    uint32_t* d_sides = nullptr;
    int side_words = 0;
    float side_iso = __builtin_nanf("");
    std::vector<int32_t> table;
    std::vector<int64_t> slot_brick;
    std::vector<uint8_t> ref;
    PinBuf<uint8_t> h_ref;   // the device reference bytes, read back after each render pass
    std::vector<uint64_t> pin;
    uint64_t epoch = 0;
    uint64_t protect_from = 0;    // bricks loaded in epochs >= this (since the last frame ended) go last
    int64_t hand = 0, used = 0;   // slots [used, gpu_slots) were never filled
    // Slots below `used` that hold no brick.  Loads issued while a render
    // pass runs take only these (or never-filled slots), since no kernel can
    // be reading them; evictions happen only while no render kernel runs.
    std::vector<int64_t> free_list;
    // sources
    const float* d_vol = nullptr;
    int vol_dims[3] = {0, 0, 0};
    std::unique_ptr<HostTier> host;
    // staging of one load: the arrays of StageLayout in one pinned buffer,
    // moved to the device with one copy
    PinBuf<uint8_t> h_stage;
    DevBuf<uint8_t> d_stage;
    // stats
    int64_t stat_missed = 0, stat_uploads = 0, stat_evictions = 0;

    ~BrickCache()
    {
        host.reset();
        if (d_pool) (void)hipFree(d_pool);
        if (d_table) (void)hipFree(d_table);
        if (d_ref) (void)hipFree(d_ref);
        if (d_submm) (void)hipFree(d_submm);
        if (d_sides) (void)hipFree(d_sides);
        h_stage.release();
        d_stage.release();
        h_ref.release();
    }

    int64_t resident() const { return used - (int64_t)free_list.size(); }
    int64_t available() const { return (int64_t)free_list.size() + (gpu_slots - used); }

    // a slot that holds no brick, or -1
    int64_t take_free()
    {
        if (!free_list.empty()) {
            int64_t s = free_list.back();
            free_list.pop_back();
            return s;
        }
        return used < gpu_slots ? used++ : -1;
    }

    // Frames in flight hold the bricks loaded for rays that have not yet run
    // the launch that uses them: slots pinned in epochs >= hold_from are
    // skipped by the clock as long as another victim exists (~0: no hold).
    uint64_t hold_from = ~0ull;

    // clock over the slots that hold bricks, skipping those pinned in the
    // current epoch (and those held, with `hold`; and, when `recent`, those
    // loaded since the last frame ended: prefetched or requested for the
    // frame in progress) and giving referenced ones a second chance
    int64_t clock_victim(bool recent, bool hold = true)
    {
        for (int64_t it = 0; it < 3 * gpu_slots; ++it) {
            int64_t c = hand;
            hand = (hand + 1) % gpu_slots;
            if (c >= used || slot_brick[(size_t)c] < 0) continue;
            const uint64_t p = pin[(size_t)c];
            if (p == epoch || (hold && p >= hold_from) || (recent && p >= protect_from)) continue;
            if (ref[(size_t)c]) { ref[(size_t)c] = 0; continue; }
            return c;
        }
        return -1;
    }

    uint64_t recent_exhausted = ~0ull;   // epoch in which every unprotected slot was taken

    // A slot for a load between passes: free, else a clock victim, held
    // slots last (a load of the current epoch always finds one, so every
    // service makes progress).
    int64_t pick_victim()
    {
        int64_t s = take_free();
        if (s < 0 && recent_exhausted != epoch) {
            s = clock_victim(true);
            if (s < 0) recent_exhausted = epoch;   // do not rescan the ring for every victim
        }
        if (s < 0) s = clock_victim(false);
        return s >= 0 || hold_from == ~0ull ? s : clock_victim(false, false);
    }
};

}  // namespace vt_rt

using vt_rt::fail;
using namespace vt_rt;

// Runtime steps shared by the single-frame loop (vt_runtime.cpp), the
// multi-frame loop (vt_frames.cpp) and the prefetcher (vt_prefetch.cpp).
namespace vt_rt {
int set_device(vt_context* ctx);
OctreeView make_view(const vt_context* ctx, int final_level);
Camera to_cam(const vt_camera* c);
int check_camera(const vt_camera* cam);
}  // namespace vt_rt
// Device reference bits into the host clock (copied: already read back).
int fold_ref_bits(vt_context* ctx, bool copied = false);
// Loads requested bricks (indices relative to 8^levelC) between passes; with
// `reset`, only enqueued, its last kernel resetting the per-pass state.
int service_requests(vt_context* ctx, const uint32_t* reqs, int64_t count, vt_frame_stats* st,
                     const PassReset* reset = nullptr);
int copy_stream_idle(vt_context* ctx);
int ensure_sides(vt_context* ctx, float iso);
int publish_loads(vt_context* ctx, hipStream_t s);
int make_reserve(vt_context* ctx, int64_t reserve);
int load_during_pass(vt_context* ctx, hipEvent_t done, int64_t* loaded, int64_t* misses);
int host_tier_start(vt_context* ctx);
struct vt_mlp;
// vt_prefetch's scoring, selection and loads; without `publish` the copies'
// sub-block min/max and page-table entries wait for publish_loads, so a
// render pass may run while they land, and only free slots are filled (no
// resident brick is evicted for a prefetch).
int prefetch_bricks(vt_context* ctx, vt_mlp* m, const vt_camera* cam, int64_t budget, float min_prob, bool publish,
                    int64_t* loaded);

struct vt_file {
    H5Volume vol;
};

struct vt_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t cstream = nullptr;   // prefetch loads (vt_render_frame waits for them)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_copy = nullptr;   // the copy stream's work so far, for publishing its loads
    // octree
    bool octree_ready = false;
    int n = 0, max_level = 0;
    uint8_t* d_child = nullptr;
    uint64_t level_off[24] = {0};
    int root_valid = 0;
    // cache
    std::unique_ptr<BrickCache> cache;
    // ray state
    DevBuf<uint64_t> ray_node;
    DevBuf<uint8_t> ray_phase;
    DevBuf<uint64_t> ray_defer;   // [max_defer][R] deferred nodes of parked rays
    DevBuf<uint32_t> list[2];
    DevBuf<HitRecord> hits_scratch;
    DevBuf<uint32_t> req_bits;
    // request list of a pass: brick index + 1 per entry, written by the
    // render kernel with system-scope stores into coherent host memory and
    // read by the host after the pass (no copy back); 0 = not written
    PinBuf<uint32_t> req_list;
    AsyncStage astage;
    std::vector<uint32_t> req_left;   // requests of a pass left for the synchronous service
    DevBuf<uint8_t> trace;
    FrameCounters* d_ctr = nullptr;
    FrameCounters* h_ctr = nullptr;
    int collect_stats = 0, record_trace = 0;
    int skip_empty = 1;   // sub-block sample skipping in the march
    int max_defer = 1;    // absent-brick nodes a ray walks past in the first pass
    int max_defer_late = 2;   // ... and in later passes
    int64_t wide_lanes = 1 << 18;   // lane budget of the wide march in later passes (0: off, -k: 2^k lanes)
    int64_t wide_first = 1 << 21;   // ... and in the first pass
    int64_t early_reserve = 256;    // free slots kept for loads issued while the first pass runs (0: off)
    int first_nodes = 16;           // walk nodes per ray in the first pass before it parks (0: no limit)
    int supertile = 4;             // first passes: S x S tile blocks per XCD (1: single tiles)
    bool pass_state_clean = false;  // the last frame left counters and request bits at zero
    int64_t early_last = INT64_MAX;  // bricks the previous frame's first pass requested
    int persistent = 0;   // persistent render pass with lane refill (measured slower; option)
    int persistent_wgs = 0;
    DevBuf<uint32_t> d_next;   // its 8 work counters
    int64_t trace_n = 0;
    DevBuf<uint8_t> stage_trav;
    int64_t stage_rays = 0;
    // host copies of the uploaded run lists (levels 0..max_level), and the
    // candidate brick codes derived from them for the prefetcher
    std::vector<std::vector<uint64_t>> up_runs;
    int cand_level = -1;
    std::vector<uint64_t> cand_codes;
    DevBuf<uint64_t> d_cand;
    // last octree build (all levels, host copies)
    std::vector<std::vector<uint64_t>> built_runs;
    float built_iso = 0.0f;
    int built_n = 0, built_dim = 0;
    int built_real[3] = {0, 0, 0};

    ~vt_context()
    {
        cache.reset();
        if (d_child) (void)hipFree(d_child);
        ray_node.release(); ray_phase.release(); ray_defer.release(); d_next.release(); list[0].release(); list[1].release();
        if (cstream) (void)hipStreamSynchronize(cstream);
        hits_scratch.release(); req_bits.release(); req_list.release(); trace.release(); stage_trav.release();
        astage.release();
        d_cand.release();
        if (d_ctr) (void)hipFree(d_ctr);
        if (h_ctr) (void)hipHostFree(h_ctr);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (ev_copy) (void)hipEventDestroy(ev_copy);
        if (stream) (void)hipStreamDestroy(stream);
        if (cstream) (void)hipStreamDestroy(cstream);
    }
};

