// vt_serve.cpp -- render-request serving loop (no reference counterpart).
//
// Scenes are volumes whose every brick is resident in HBM, with a min/max
// pyramid over the octree levels so that any isovalue gets its membership
// tables without rebuilding the octree (a node is valid for iso t exactly
// when min < t <= max over its zero-padded samples, the closed form of the
// reference builder).
//
// A worker thread serves requests as they are submitted (continuous
// batching): it admits a request while its outputs (48 B per ray) and its
// isovalue's membership tables fit one HBM arena of the server's budget,
// and every launch renders 16x16 tiles of all admitted requests together.
// Two batches are in flight at once (double-buffered tile lists), so the
// host admits and prepares the next batch while the device renders the
// current one.  A launch is a captured hipGraph, replayed per batch with a
// new tile list, in a few grid sizes so that a small batch does not dispatch
// the largest grid.  Mode 1 renders one request per launch, in arrival
// order (the plain per-request baseline).
#include <math.h>

#include <deque>
#include <map>
#include <thread>
#include <unordered_map>

#include "vt_build.hpp"
#include "vt_context.hpp"

namespace {

struct Scene {
    int n = 0, level_o = 0, level_c = 0, inc = 0, S = 0;
    size_t slot_elems = 0;
    float* d_pool = nullptr;
    int32_t* d_slot = nullptr;
    float2* d_submm = nullptr;
    int G = 0;
    std::vector<std::vector<float>> mn, mx;   // per level 0..levelO, 8^l nodes
    struct Table {
        size_t off = 0;      // membership bytes in the server's arena
        int root = 0;        // the root node is valid
        int refs = 0;        // admitted requests using it
        uint64_t used = 0;   // last use (LRU among unreferenced tables)
    };
    std::map<uint32_t, Table> tables;         // isovalue bits -> membership tables
    std::vector<uint64_t> level_off;
    size_t table_bytes = 0;
    float iso_lo = -INFINITY, iso_hi = INFINITY;   // isovalues the resident bricks serve
    int64_t resident = 0;
};

struct Arena {   // first-fit allocator over one device allocation
    char* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_;   // offset -> length
    bool init(size_t bytes)
    {
        if (hipMalloc((void**)&base, bytes) != hipSuccess) return false;
        size = bytes;
        free_[0] = bytes;
        return true;
    }
    size_t alloc(size_t bytes)
    {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes) {
                size_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > bytes) free_[off + bytes] = len - bytes;
                return off;
            }
        return (size_t)-1;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = (bytes + 255) & ~(size_t)255;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

constexpr int kBufs = 2;        // batches in flight (double-buffered tile lists)
constexpr int kBuckets = 7;     // captured grid sizes: max_tiles / 4^(kBuckets-1-b) tiles

struct Pending {
    int64_t ticket;
    vt_request q;
};

struct Ticket {
    int state = 0;               // 0 queued, 1 admitted, 2 done
    int rc = VT_OK;
    std::string msg;             // the failure's message (set on the worker's thread)
    double t_submit = 0, t_admit = 0, t_done = 0;   // host clock, ms
    int32_t launches = 0;
};

struct Job {                     // an admitted request
    int64_t ticket = 0;
    vt_request q;
    int slot = 0;                // entry of the job table on the device
    size_t off = 0, bytes = 0;   // outputs in the arena
    Scene* sc = nullptr;
    uint32_t key = 0;            // its membership table
    int tiles = 0, next_tile = 0, launches = 0;
    int64_t last_batch = -1;     // batch that holds its last tile
    int rc = VT_OK;
};

struct Batch {
    int64_t seq = 0;
    int buf = 0;
    std::vector<int64_t> tickets;   // jobs with tiles in this batch
};

}  // namespace

struct vt_server {
    vt_context* ctx = nullptr;
    hipStream_t stream = nullptr;    // batches
    hipStream_t ostream = nullptr;   // output copies to host memory
    int64_t budget = 0;
    int max_tiles = 0;
    int max_jobs = 256;
    std::vector<std::unique_ptr<Scene>> scenes;
    Arena arena;
    DevBuf<ServeJob> d_jobs;
    PinBuf<ServeJob> h_jobs;   // pinned staging, one entry per job slot (async copies read it later)
    DevBuf<ServeTile> d_work[kBufs];
    PinBuf<ServeTile> h_work[kBufs];
    ServeCounters* d_ctr[kBufs] = {};
    PinBuf<ServeCounters> h_ctr[kBufs];   // counters uploaded with a batch
    PinBuf<ServeCounters> h_res[kBufs];   // counters read back after it
    hipEvent_t ev[kBufs] = {};
    // one captured launch per lane width 2^lg, grid bucket and buffer
    static constexpr int kServeWidths = 4;
    hipGraph_t graph[kServeWidths][kBuckets][kBufs] = {};
    hipGraphExec_t exec[kServeWidths][kBuckets][kBufs] = {};
    int bucket_tiles[kBuckets] = {};
    int64_t lane_budget = 1 << 21;   // lanes of a batch (rays x 2^lg); <= 0: fixed 2^-lane_budget lanes
    // staging of membership tables (rewritten once the last copy from it ended)
    PinBuf<uint8_t> h_table;
    hipEvent_t ev_table = nullptr;
    uint64_t table_clock = 0;
    int64_t tables_built = 0;
    // online state
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Pending> queue;
    std::unordered_map<int64_t, Ticket> tickets;
    int64_t next_ticket = 1;
    std::thread worker;
    bool running = false, stopping = false;
    int mode = 0;
    int worker_rc = VT_OK;
    std::string worker_err;
    // statistics since vt_server_start
    int64_t st_requests = 0, st_launches = 0, st_tiles = 0, st_rays = 0;
    double st_busy_ms = 0.0, st_t0 = 0.0, st_t1 = 0.0;

    ~vt_server()
    {
        for (int w = 0; w < kServeWidths; ++w)
            for (int b = 0; b < kBuckets; ++b)
                for (int k = 0; k < kBufs; ++k) {
                    if (exec[w][b][k]) (void)hipGraphExecDestroy(exec[w][b][k]);
                    if (graph[w][b][k]) (void)hipGraphDestroy(graph[w][b][k]);
                }
        for (auto& s : scenes) {
            if (s->d_pool) (void)hipFree(s->d_pool);
            if (s->d_slot) (void)hipFree(s->d_slot);
            if (s->d_submm) (void)hipFree(s->d_submm);
        }
        if (arena.base) (void)hipFree(arena.base);
        for (int k = 0; k < kBufs; ++k) {
            if (d_ctr[k]) (void)hipFree(d_ctr[k]);
            if (ev[k]) (void)hipEventDestroy(ev[k]);
            d_work[k].release();
            h_work[k].release();
            h_ctr[k].release();
            h_res[k].release();
        }
        if (ev_table) (void)hipEventDestroy(ev_table);
        d_jobs.release();
        h_jobs.release();
        h_table.release();
        if (stream) (void)hipStreamDestroy(stream);
        if (ostream) (void)hipStreamDestroy(ostream);
    }
};

namespace {

int tree_levels3(int X, int Y, int Z)
{
    int d = std::max(X, std::max(Y, Z)), n = 0;
    while ((1 << n) < d) ++n;
    return n;
}

// Frees the least recently used membership table that no admitted request
// uses; false when there is none.
bool evict_table(vt_server* sv)
{
    Scene* best_s = nullptr;
    uint32_t best_k = 0;
    uint64_t best_u = ~0ull;
    for (auto& s : sv->scenes)
        for (auto& t : s->tables)
            if (t.second.refs == 0 && t.second.used < best_u) {
                best_u = t.second.used;
                best_s = s.get();
                best_k = t.first;
            }
    if (!best_s) return false;
    sv->arena.release(best_s->tables[best_k].off, best_s->table_bytes);
    best_s->tables.erase(best_k);
    return true;
}

// Arena space for `bytes`, evicting unused membership tables as needed;
// (size_t)-1 when it does not fit now.
size_t arena_take(vt_server* sv, size_t bytes)
{
    for (;;) {
        size_t off = sv->arena.alloc(bytes);
        if (off != (size_t)-1) return off;
        if (!evict_table(sv)) return (size_t)-1;
    }
}

// The membership tables of levels 1..levelO (one byte per parent) for
// isovalue `iso`, in the arena, with one more reference.  VT_OK and *table
// set; 1 when the arena has no room now; a negative status on failure.
int acquire_table(vt_server* sv, Scene& sc, float iso, Scene::Table** table)
{
    uint32_t key;
    memcpy(&key, &iso, 4);
    auto it = sc.tables.find(key);
    if (it == sc.tables.end()) {
        const size_t bytes = std::max<size_t>(1, sc.table_bytes);
        const size_t off = arena_take(sv, bytes);
        if (off == (size_t)-1) return 1;
        // the staging may still feed the previous table's copy
        HIP_TRY(hipEventSynchronize(sv->ev_table));
        if (sv->h_table.ensure(bytes) != hipSuccess) {
            sv->arena.release(off, bytes);
            return fail(VT_E_NOMEM, "table staging");
        }
        uint8_t* h = sv->h_table.p;
        memset(h, 0, bytes);
        for (int l = 1; l <= sc.level_o; ++l) {
            const auto& mn = sc.mn[(size_t)l];
            const auto& mx = sc.mx[(size_t)l];
            for (size_t c = 0; c < mn.size(); ++c)
                if (mn[c] < iso && iso <= mx[c]) h[(size_t)(sc.level_off[(size_t)l] + (c >> 3))] |= (uint8_t)(1u << (c & 7));
        }
        // ahead of every batch that uses it, on the batch stream
        HIP_TRY(hipMemcpyAsync(sv->arena.base + off, h, bytes, hipMemcpyHostToDevice, sv->stream));
        HIP_TRY(hipEventRecord(sv->ev_table, sv->stream));
        Scene::Table t;
        t.off = off;
        t.root = sc.mn[0][0] < iso && iso <= sc.mx[0][0];
        it = sc.tables.emplace(key, t).first;
        sv->tables_built++;
    }
    it->second.refs++;
    it->second.used = ++sv->table_clock;
    *table = &it->second;
    return VT_OK;
}

void release_table(Scene& sc, uint32_t key)
{
    auto it = sc.tables.find(key);
    if (it != sc.tables.end() && it->second.refs > 0) it->second.refs--;
}

// Marks a request done (under the lock) and wakes its poller.
void complete(vt_server* sv, int64_t ticket, int rc, int32_t launches)
{
    std::lock_guard<std::mutex> g(sv->m);
    Ticket& t = sv->tickets[ticket];
    t.state = 2;
    t.rc = rc;
    if (rc != VT_OK) t.msg = vt_rt::g_err;
    t.t_done = now_ms();
    t.launches = launches;
    sv->cv_done.notify_all();
}

// Admits `p` into job slot `slot`: 0 admitted, 1 no room now, negative: the
// request failed (its status).
int admit(vt_server* sv, const Pending& p, int slot, Job& J)
{
    const vt_request& q = p.q;
    Scene& sc = *sv->scenes[(size_t)q.scene];
    const size_t R = (size_t)q.cam.W * q.cam.H;
    const size_t bytes = R * 48;
    Scene::Table* table = nullptr;
    int rc = acquire_table(sv, sc, q.iso, &table);
    if (rc != VT_OK) return rc;
    uint32_t key;
    memcpy(&key, &q.iso, 4);
    const size_t off = arena_take(sv, bytes);
    if (off == (size_t)-1) {
        release_table(sc, key);
        return 1;
    }
    ServeJob& S = sv->h_jobs.p[slot];
    memcpy(&S.cam, &q.cam, sizeof S.cam);
    S.oct.child_bytes = reinterpret_cast<const uint8_t*>(sv->arena.base + table->off);
    for (int l = 0; l < 24; ++l) S.oct.level_off[l] = sc.level_off[(size_t)l];
    S.oct.n = sc.n;
    S.oct.final_level = sc.level_o;
    S.oct.root_valid = table->root;
    S.bv = BrickView{sc.d_slot, sc.d_pool, sc.S, sc.inc, sc.slot_elems, sc.level_c,
                     3 * (sc.level_o - sc.level_c), 1ull << (3 * sc.level_c), sc.d_submm, sc.G};
    S.sh = Shading{q.iso, q.light[0], q.light[1], q.light[2], q.color_scale};
    S.rgba = reinterpret_cast<float4*>(sv->arena.base + off);
    S.hits = reinterpret_cast<HitRecord*>(sv->arena.base + off + R * 16);
    S.tiles_x = (q.cam.W + 15) / 16;
    // the slot's staging is rewritten only after its previous job retired,
    // i.e. after every batch that read it
    if (hipMemcpyAsync(sv->d_jobs.p + slot, &S, sizeof S, hipMemcpyHostToDevice, sv->stream) != hipSuccess) {
        sv->arena.release(off, bytes);
        release_table(sc, key);
        return fail(VT_E_HIP, "job upload failed");
    }
    J = Job();
    J.ticket = p.ticket;
    J.q = q;
    J.slot = slot;
    J.off = off;
    J.bytes = bytes;
    J.sc = &sc;
    J.key = key;
    J.tiles = S.tiles_x * ((q.cam.H + 15) / 16);
    return VT_OK;
}

// Copies a finished request's outputs to its host buffers (if any), frees
// its arena space and table reference.
int retire(vt_server* sv, Job& J)
{
    const size_t R = (size_t)J.q.cam.W * J.q.cam.H;
    int rc = J.rc;
    if (rc == VT_OK && (J.q.h_rgba || J.q.h_hits)) {
        hipError_t e = hipSuccess;
        if (J.q.h_rgba)
            e = hipMemcpyAsync(J.q.h_rgba, sv->arena.base + J.off, R * 16, hipMemcpyDeviceToHost, sv->ostream);
        if (e == hipSuccess && J.q.h_hits)
            e = hipMemcpyAsync(J.q.h_hits, sv->arena.base + J.off + R * 16, R * 32, hipMemcpyDeviceToHost,
                               sv->ostream);
        if (e == hipSuccess) e = hipStreamSynchronize(sv->ostream);
        if (e != hipSuccess) rc = fail(VT_E_HIP, std::string("output copy: ") + hipGetErrorString(e));
    }
    sv->arena.release(J.off, J.bytes);
    release_table(*J.sc, J.key);
    sv->st_rays += (int64_t)R;
    complete(sv, J.ticket, rc, J.launches);
    return rc;
}

// The worker: admit, batch, launch, retire, until stopped with nothing left.
void serve_loop(vt_server* sv)
{
    int fatal = VT_OK;
    if (hipSetDevice(sv->ctx->device) != hipSuccess) fatal = fail(VT_E_HIP, "hipSetDevice");
    std::deque<Pending> pending;
    std::vector<Job> live;
    std::vector<int> free_slots;
    for (int j = sv->max_jobs - 1; j >= 0; --j) free_slots.push_back(j);
    std::deque<Batch> inflight;
    int64_t seq = 0;
    double busy_t0 = 0.0;
    auto fail_all = [&](int rc) {
        for (auto& J : live) {
            J.rc = rc;
            (void)retire(sv, J);
        }
        live.clear();
        for (auto& p : pending) complete(sv, p.ticket, rc, 0);
        pending.clear();
        std::lock_guard<std::mutex> g(sv->m);
        for (auto& p : sv->queue) {
            Ticket& t = sv->tickets[p.ticket];
            t.state = 2;
            t.rc = rc;
            t.msg = sv->worker_err;
            t.t_done = now_ms();
        }
        sv->queue.clear();
        sv->cv_done.notify_all();
    };
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(sv->m);
            if (fatal == VT_OK && pending.empty() && live.empty() && inflight.empty())
                sv->cv_work.wait(lk, [&] { return sv->stopping || !sv->queue.empty(); });
            while (!sv->queue.empty()) {
                pending.push_back(sv->queue.front());
                sv->queue.pop_front();
            }
            if (fatal != VT_OK) {
                sv->worker_rc = fatal;
                sv->worker_err = vt_rt::g_err;
                lk.unlock();
                fail_all(fatal);
                lk.lock();
                if (sv->stopping) break;
                sv->cv_work.wait(lk, [&] { return sv->stopping || !sv->queue.empty(); });
                continue;
            }
            if (sv->stopping && pending.empty() && live.empty() && inflight.empty()) break;
        }
        // admit in submission order while job slots and the arena allow (mode 1:
        // one request at a time, on an idle device)
        while (!pending.empty() && !free_slots.empty() && (sv->mode == 0 || (live.empty() && inflight.empty()))) {
            Job J;
            int rc = admit(sv, pending.front(), free_slots.back(), J);
            if (rc == 1 && live.empty() && inflight.empty())
                rc = fail(VT_E_NOMEM, "a request's outputs and tables do not fit the HBM budget");
            if (rc == 1) break;   // wait for outputs to drain
            const int64_t ticket = pending.front().ticket;
            pending.pop_front();
            if (rc != VT_OK) {
                complete(sv, ticket, rc, 0);
                continue;
            }
            free_slots.pop_back();
            {
                std::lock_guard<std::mutex> g(sv->m);
                Ticket& t = sv->tickets[ticket];
                t.state = 1;
                t.t_admit = now_ms();
            }
            live.push_back(J);
        }
        // fill and launch batches while a buffer is free, oldest request first
        while ((int)inflight.size() < kBufs && (sv->mode == 0 || inflight.empty())) {
            const int b = (int)(seq % kBufs);
            Batch B;
            B.seq = seq;
            B.buf = b;
            int nw = 0;
            for (auto& J : live) {
                if (J.next_tile >= J.tiles || nw >= sv->max_tiles) continue;
                while (J.next_tile < J.tiles && nw < sv->max_tiles) {
                    sv->h_work[b].p[nw].job = (uint32_t)J.slot;
                    sv->h_work[b].p[nw].tile = (uint32_t)J.next_tile++;
                    ++nw;
                }
                J.launches++;
                B.tickets.push_back(J.ticket);
                if (J.next_tile == J.tiles) J.last_batch = seq;
            }
            if (nw == 0) break;
            sv->h_ctr[b].p->nwork = (uint32_t)nw;
            sv->h_ctr[b].p->err = 0;
            int lg = 0;
            if (sv->lane_budget > 0)
                while (lg + 1 < vt_server::kServeWidths && ((int64_t)nw * 256 << (lg + 1)) <= sv->lane_budget) ++lg;
            else
                lg = (int)std::min<int64_t>(-sv->lane_budget, vt_server::kServeWidths - 1);
            int bk = 0;
            while (bk + 1 < kBuckets && sv->bucket_tiles[bk] < nw) ++bk;
            hipError_t e = hipMemcpyAsync(sv->d_work[b].p, sv->h_work[b].p, (size_t)nw * sizeof(ServeTile),
                                          hipMemcpyHostToDevice, sv->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(sv->d_ctr[b], sv->h_ctr[b].p, sizeof(ServeCounters), hipMemcpyHostToDevice,
                                   sv->stream);
            if (e == hipSuccess) e = hipGraphLaunch(sv->exec[lg][bk][b], sv->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(sv->h_res[b].p, sv->d_ctr[b], sizeof(ServeCounters), hipMemcpyDeviceToHost,
                                   sv->stream);
            if (e == hipSuccess) e = hipEventRecord(sv->ev[b], sv->stream);
            if (e != hipSuccess) {
                fatal = fail(VT_E_HIP, std::string("serving launch: ") + hipGetErrorString(e));
                break;
            }
            if (inflight.empty()) busy_t0 = now_ms();
            inflight.push_back(std::move(B));
            sv->st_launches++;
            sv->st_tiles += nw;
            ++seq;
        }
        if (fatal != VT_OK) {
            (void)hipStreamSynchronize(sv->stream);
            inflight.clear();
            continue;
        }
        // wait for the oldest batch and retire the requests it finished
        if (!inflight.empty()) {
            Batch B = std::move(inflight.front());
            inflight.pop_front();
            hipError_t e = hipEventSynchronize(sv->ev[B.buf]);
            if (inflight.empty()) sv->st_busy_ms += now_ms() - busy_t0;
            if (e != hipSuccess) {
                fatal = fail(VT_E_HIP, std::string("serving batch: ") + hipGetErrorString(e));
                continue;
            }
            const uint32_t err = sv->h_res[B.buf].p->err;
            for (size_t k = 0; k < live.size();) {
                Job& J = live[k];
                if (err && std::find(B.tickets.begin(), B.tickets.end(), J.ticket) != B.tickets.end() &&
                    J.rc == VT_OK)
                    J.rc = fail(VT_E_RANGE, "serving pass: a sample fell outside its brick slot");
                if (J.last_batch != B.seq) {
                    ++k;
                    continue;
                }
                (void)retire(sv, J);
                free_slots.push_back(J.slot);
                live.erase(live.begin() + (long)k);
            }
        }
    }
}

int worker_start(vt_server* sv, int mode)
{
    std::lock_guard<std::mutex> g(sv->m);
    if (sv->running) return fail(VT_E_STATE, "the server is already running");
    sv->mode = mode;
    sv->stopping = false;
    sv->worker_rc = VT_OK;
    sv->st_requests = sv->st_launches = sv->st_tiles = sv->st_rays = 0;
    sv->st_busy_ms = 0.0;
    sv->st_t0 = now_ms();
    sv->running = true;
    sv->worker = std::thread(serve_loop, sv);
    return VT_OK;
}

int worker_stop(vt_server* sv, vt_serve_stats* stats)
{
    {
        std::lock_guard<std::mutex> g(sv->m);
        if (!sv->running) return fail(VT_E_STATE, "the server is not running");
        sv->stopping = true;
        sv->cv_work.notify_all();
    }
    sv->worker.join();
    std::lock_guard<std::mutex> g(sv->m);
    sv->running = false;
    sv->st_t1 = now_ms();
    if (sv->worker_rc != VT_OK) (void)fail(sv->worker_rc, sv->worker_err);
    if (stats) {
        stats->requests = sv->st_requests;
        stats->launches = sv->st_launches;
        stats->tiles = sv->st_tiles;
        stats->rays = sv->st_rays;
        stats->wall_s = (sv->st_t1 - sv->st_t0) * 1e-3;
        stats->busy_s = sv->st_busy_ms * 1e-3;
        stats->tables_built = sv->tables_built;
        int64_t held = 0;
        for (auto& s : sv->scenes) held += (int64_t)s->tables.size();
        stats->tables_held = held;
    }
    return sv->worker_rc;
}

int check_request(vt_server* sv, const vt_request& q)
{
    if (q.scene < 0 || q.scene >= (int32_t)sv->scenes.size()) return fail(VT_E_ARG, "unknown scene");
    const Scene& qs = *sv->scenes[(size_t)q.scene];
    if (!(q.iso >= qs.iso_lo && q.iso <= qs.iso_hi)) return fail(VT_E_ARG, "isovalue outside the scene's range");
    if (q.cam.W <= 0 || q.cam.H <= 0 || (int64_t)q.cam.W * q.cam.H > (1 << 26)) return fail(VT_E_ARG, "bad screen");
    const size_t need = (size_t)q.cam.W * q.cam.H * 48 + ((std::max<size_t>(1, qs.table_bytes) + 255) & ~(size_t)255);
    if ((int64_t)need > sv->budget) return fail(VT_E_ARG, "a request's outputs and tables exceed the HBM budget");
    return VT_OK;
}

int submit(vt_server* sv, const vt_request& q, int64_t* ticket)
{
    if (int rc = check_request(sv, q)) return rc;
    std::lock_guard<std::mutex> g(sv->m);
    if (!sv->running) return fail(VT_E_STATE, "the server is not running (vt_server_start)");
    if (sv->worker_rc != VT_OK) return fail(sv->worker_rc, "the serving worker failed: " + sv->worker_err);
    const int64_t t = sv->next_ticket++;
    Ticket& k = sv->tickets[t];
    k.t_submit = now_ms();
    sv->queue.push_back(Pending{t, q});
    sv->st_requests++;
    sv->cv_work.notify_one();
    *ticket = t;
    return VT_OK;
}

// done: the request's status, its ticket forgotten; pending: *done = 0
int poll(vt_server* sv, int64_t ticket, int wait, int32_t* done, Ticket* out)
{
    std::unique_lock<std::mutex> lk(sv->m);
    auto it = sv->tickets.find(ticket);
    if (it == sv->tickets.end()) return fail(VT_E_ARG, "unknown ticket");
    if (wait) sv->cv_done.wait(lk, [&] { return sv->tickets[ticket].state == 2; });
    it = sv->tickets.find(ticket);
    if (it->second.state != 2) {
        *done = 0;
        return VT_OK;
    }
    *done = 1;
    *out = it->second;
    sv->tickets.erase(it);
    // the message of a failure on the worker's thread, on the caller's
    return out->rc == VT_OK ? VT_OK : fail(out->rc, out->msg);
}

}  // namespace

extern "C" {

int vt_server_create(vt_context* ctx, int64_t hbm_budget_bytes, int32_t max_tiles, vt_server** out)
{
    if (!ctx || !out) return fail(VT_E_ARG, "NULL argument");
    if (hbm_budget_bytes < (1 << 20) || max_tiles < 1 || max_tiles > (1 << 20)) return fail(VT_E_ARG, "bad sizes");
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<vt_server> sv(new vt_server);
    sv->ctx = ctx;
    sv->budget = hbm_budget_bytes;
    sv->max_tiles = max_tiles;
    HIP_TRY(hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&sv->ostream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&sv->ev_table, hipEventDisableTiming));
    if (!sv->arena.init((size_t)hbm_budget_bytes)) return fail(VT_E_NOMEM, "serving arena");
    if (sv->d_jobs.ensure((size_t)sv->max_jobs) != hipSuccess || sv->h_jobs.ensure((size_t)sv->max_jobs) != hipSuccess)
        return fail(VT_E_NOMEM, "serving buffers");
    for (int k = 0; k < kBufs; ++k) {
        if (sv->d_work[k].ensure((size_t)max_tiles) != hipSuccess || sv->h_work[k].ensure((size_t)max_tiles) != hipSuccess ||
            sv->h_ctr[k].ensure(1) != hipSuccess || sv->h_res[k].ensure(1) != hipSuccess)
            return fail(VT_E_NOMEM, "serving buffers");
        HIP_TRY(hipMalloc((void**)&sv->d_ctr[k], sizeof(ServeCounters)));
        HIP_TRY(hipMemset(sv->d_ctr[k], 0, sizeof(ServeCounters)));
        HIP_TRY(hipEventCreateWithFlags(&sv->ev[k], hipEventDisableTiming));
    }
    for (int b = 0; b < kBuckets; ++b)
        sv->bucket_tiles[b] = std::max(1, max_tiles >> (2 * (kBuckets - 1 - b)));
    // the batch launch per lane width, grid size and buffer, captured once and
    // replayed: the tile list and count change between replays, the shape never does
    for (int lg = 0; lg < vt_server::kServeWidths; ++lg)
        for (int b = 0; b < kBuckets; ++b)
            for (int k = 0; k < kBufs; ++k) {
                HIP_TRY(hipStreamBeginCapture(sv->stream, hipStreamCaptureModeThreadLocal));
                hipError_t le = launch_serve_pass(sv->d_jobs.p, sv->d_work[k].p, sv->d_ctr[k], sv->bucket_tiles[b],
                                                  lg, sv->stream);
                hipError_t ce = hipStreamEndCapture(sv->stream, &sv->graph[lg][b][k]);
                if (le != hipSuccess) return fail(VT_E_HIP, std::string("capture: ") + hipGetErrorString(le));
                if (ce != hipSuccess) return fail(VT_E_HIP, std::string("capture: ") + hipGetErrorString(ce));
                HIP_TRY(hipGraphInstantiate(&sv->exec[lg][b][k], sv->graph[lg][b][k], nullptr, nullptr, 0));
            }
    *out = sv.release();
    return VT_OK;
}

int vt_server_set_lanes(vt_server* sv, int64_t lane_budget)
{
    if (!sv) return fail(VT_E_ARG, "sv is NULL");
    if (lane_budget < -(vt_server::kServeWidths - 1)) return fail(VT_E_ARG, "lane_budget below the widest width");
    sv->lane_budget = lane_budget;
    return VT_OK;
}

int vt_server_destroy(vt_server* sv)
{
    if (sv) {
        bool run;
        {
            std::lock_guard<std::mutex> g(sv->m);
            run = sv->running;
        }
        if (run) (void)worker_stop(sv, nullptr);
        (void)hipStreamSynchronize(sv->stream);
        (void)hipStreamSynchronize(sv->ostream);
        delete sv;
    }
    return VT_OK;
}

// A scene serves requests with isovalues in [iso_lo, iso_hi].  A request at
// iso visits only nodes with min < iso <= max (the membership tables below),
// so every brick it marches has min < iso_hi and max >= iso_lo at levelC:
// those bricks are the resident ones (all of them for the full range).
static int add_scene(vt_server* sv, const float* d_vol, int32_t X, int32_t Y, int32_t Z, int32_t levelO,
                     int32_t levelC, int32_t cubeInc, float iso_lo, float iso_hi, int32_t* scene_out,
                     int64_t* resident_out)
{
    if (!sv || !d_vol || !scene_out) return fail(VT_E_ARG, "NULL argument");
    {
        std::lock_guard<std::mutex> g(sv->m);
        if (sv->running) return fail(VT_E_STATE, "scenes are added while the server is stopped");
    }
    if (!(iso_lo <= iso_hi)) return fail(VT_E_ARG, "the isovalue range must have iso_lo <= iso_hi");
    if (X <= 0 || Y <= 0 || Z <= 0) return fail(VT_E_ARG, "bad volume dims");
    const int n = tree_levels3(X, Y, Z);
    if (levelC < 0 || levelC > levelO || levelO > n || levelO > 8 || cubeInc < 0)
        return fail(VT_E_ARG, "bad levels");
    hipStream_t s = sv->ctx->stream;
    std::unique_ptr<Scene> sc(new Scene);
    sc->n = n;
    sc->level_o = levelO;
    sc->level_c = levelC;
    sc->inc = cubeInc;
    sc->iso_lo = iso_lo;
    sc->iso_hi = iso_hi;
    const int cube = 1 << (n - levelC);
    sc->S = cube + 2 * cubeInc;
    sc->slot_elems = (size_t)sc->S * sc->S * sc->S;
    // min/max of every level-levelO node on the device, then up the pyramid
    const size_t NO = (size_t)1 << (3 * levelO);
    DevBuf<float> dmn, dmx;
    if (dmn.ensure(NO) != hipSuccess || dmx.ensure(NO) != hipSuccess) return fail(VT_E_NOMEM, "pyramid");
    HIP_TRY(launch_node_minmax(d_vol, X, Y, Z, n, levelO, dmn.p, dmx.p, s));
    sc->mn.assign((size_t)levelO + 1, {});
    sc->mx.assign((size_t)levelO + 1, {});
    sc->mn[(size_t)levelO].resize(NO);
    sc->mx[(size_t)levelO].resize(NO);
    HIP_TRY(hipMemcpyAsync(sc->mn[(size_t)levelO].data(), dmn.p, NO * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(sc->mx[(size_t)levelO].data(), dmx.p, NO * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int l = levelO - 1; l >= 0; --l) {
        size_t N = (size_t)1 << (3 * l);
        sc->mn[(size_t)l].assign(N, 3.0e38f);
        sc->mx[(size_t)l].assign(N, -3.0e38f);
        for (size_t c = 0; c < N * 8; ++c) {
            sc->mn[(size_t)l][c >> 3] = std::min(sc->mn[(size_t)l][c >> 3], sc->mn[(size_t)l + 1][c]);
            sc->mx[(size_t)l][c >> 3] = std::max(sc->mx[(size_t)l][c >> 3], sc->mx[(size_t)l + 1][c]);
        }
    }
    sc->level_off.assign(24, 0);
    uint64_t off = 0;
    for (int l = 1; l <= levelO; ++l) {
        sc->level_off[(size_t)l] = off;
        off += 1ull << (3 * (l - 1));
    }
    sc->table_bytes = (size_t)off;
    // the resident bricks, in brick order, take slots 0, 1, ...; the others
    // map to -1
    const int64_t nb = 1ll << (3 * levelC);
    const auto& bmn = sc->mn[(size_t)levelC];
    const auto& bmx = sc->mx[(size_t)levelC];
    std::vector<int32_t> slot_of((size_t)nb, -1);
    std::vector<int64_t> codes;
    std::vector<int32_t> slots;
    for (int64_t b = 0; b < nb; ++b)
        if (bmn[(size_t)b] < iso_hi && bmx[(size_t)b] >= iso_lo) {
            slot_of[(size_t)b] = (int32_t)codes.size();
            slots.push_back((int32_t)codes.size());
            codes.push_back(b + nb);
        }
    const int64_t nr = (int64_t)codes.size();
    sc->resident = nr;
    size_t pool_bytes = (size_t)std::max<int64_t>(nr, 1) * sc->slot_elems * 4;
    if (hipMalloc((void**)&sc->d_pool, pool_bytes) != hipSuccess)
        return fail(VT_E_NOMEM, "scene brick pool (" + std::to_string(pool_bytes) + " bytes)");
    HIP_TRY(hipMalloc((void**)&sc->d_slot, (size_t)nb * 4));
    HIP_TRY(hipMemcpy(sc->d_slot, slot_of.data(), (size_t)nb * 4, hipMemcpyHostToDevice));
    if (cube % 4 == 0 && cubeInc >= 1) {
        sc->G = cube / 4;
        size_t mm = (size_t)std::max<int64_t>(nr, 1) * sc->G * sc->G * sc->G * sizeof(float2);
        if (hipMalloc((void**)&sc->d_submm, mm) != hipSuccess) return fail(VT_E_NOMEM, "sub-block min/max");
    }
    if (nr > 0) {
        DevBuf<int64_t> dcodes;
        DevBuf<int32_t> dslots;
        if (dcodes.ensure((size_t)nr) != hipSuccess || dslots.ensure((size_t)nr) != hipSuccess)
            return fail(VT_E_NOMEM, "codes");
        HIP_TRY(hipMemcpy(dcodes.p, codes.data(), (size_t)nr * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(dslots.p, slots.data(), (size_t)nr * 4, hipMemcpyHostToDevice));
        for (int64_t o = 0; o < nr; o += 65536) {
            int cnt = (int)std::min<int64_t>(65536, nr - o);
            HIP_TRY(launch_bricks_from_volume(d_vol, X, Y, Z, dcodes.p + o, dslots.p + o, cnt, levelC, n, cube,
                                              cubeInc, sc->slot_elems, sc->d_pool, s));
            if (sc->d_submm)
                HIP_TRY(launch_submm(sc->d_pool, dslots.p + o, cnt, sc->S, cubeInc, sc->G, sc->slot_elems,
                                     sc->d_submm, s));
        }
        HIP_TRY(hipStreamSynchronize(s));
    }
    *scene_out = (int32_t)sv->scenes.size();
    if (resident_out) *resident_out = nr;
    sv->scenes.push_back(std::move(sc));
    return VT_OK;
}

int vt_server_add_scene(vt_server* sv, const float* d_vol, int32_t X, int32_t Y, int32_t Z, int32_t levelO,
                        int32_t levelC, int32_t cubeInc, int32_t* scene_out)
{
    return add_scene(sv, d_vol, X, Y, Z, levelO, levelC, cubeInc, -INFINITY, INFINITY, scene_out, nullptr);
}

int vt_server_add_scene_range(vt_server* sv, const float* d_vol, int32_t X, int32_t Y, int32_t Z, int32_t levelO,
                              int32_t levelC, int32_t cubeInc, float iso_lo, float iso_hi, int32_t* scene_out,
                              int64_t* resident_out)
{
    return add_scene(sv, d_vol, X, Y, Z, levelO, levelC, cubeInc, iso_lo, iso_hi, scene_out, resident_out);
}

int vt_server_start(vt_server* sv, int32_t mode)
{
    if (!sv) return fail(VT_E_ARG, "sv is NULL");
    if (mode != 0 && mode != 1) return fail(VT_E_ARG, "mode must be 0 (batched) or 1 (one request per launch)");
    return worker_start(sv, mode);
}

int vt_server_stop(vt_server* sv, vt_serve_stats* stats)
{
    if (!sv) return fail(VT_E_ARG, "sv is NULL");
    return worker_stop(sv, stats);
}

int vt_server_submit(vt_server* sv, const vt_request* req, int64_t* ticket)
{
    if (!sv || !req || !ticket) return fail(VT_E_ARG, "NULL argument");
    return submit(sv, *req, ticket);
}

int vt_server_poll(vt_server* sv, int64_t ticket, int32_t wait, int32_t* done, vt_request_result* result)
{
    if (!sv || !done) return fail(VT_E_ARG, "NULL argument");
    Ticket t;
    const int rc = poll(sv, ticket, wait, done, &t);
    if (*done && result) {
        result->admit_s = (t.t_admit > 0 ? t.t_admit - t.t_submit : t.t_done - t.t_submit) * 1e-3;
        result->done_s = (t.t_done - t.t_submit) * 1e-3;
        result->latency_ms = t.t_done - t.t_submit;
        result->launches = t.launches;
    }
    return rc;
}

// The replay form: requests submitted at their arrival times from this
// thread, served by the worker, and waited for.  Times in `results` are
// seconds after the run starts; the first failing request's status is
// returned once every request has finished.
int vt_server_run(vt_server* sv, const vt_request* reqs, int64_t nreq, int32_t mode, vt_request_result* results,
                  vt_serve_stats* stats)
{
    if (!sv || (nreq && (!reqs || !results))) return fail(VT_E_ARG, "NULL argument");
    if (mode != 0 && mode != 1) return fail(VT_E_ARG, "mode must be 0 (batched) or 1 (one request per launch)");
    for (int64_t i = 0; i < nreq; ++i)
        if (int rc = check_request(sv, reqs[i])) return rc;
    std::vector<int64_t> order((size_t)nreq);
    for (int64_t i = 0; i < nreq; ++i) order[(size_t)i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    if (int rc = worker_start(sv, mode)) return rc;
    const double t0 = now_ms();
    std::vector<int64_t> tickets((size_t)nreq, 0);
    int first_rc = VT_OK;
    for (int64_t k = 0; k < nreq; ++k) {
        const int64_t i = order[(size_t)k];
        const double wait_ms = reqs[i].arrival_s * 1e3 - (now_ms() - t0);
        if (wait_ms > 0) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(wait_ms * 1e3)));
        int rc = submit(sv, reqs[i], &tickets[(size_t)i]);
        if (rc != VT_OK) {
            first_rc = rc;
            nreq = k;   // stop submitting; wait for what was submitted
            break;
        }
    }
    for (int64_t k = 0; k < nreq; ++k) {
        const int64_t i = order[(size_t)k];
        int32_t done = 0;
        Ticket t;
        int rc = poll(sv, tickets[(size_t)i], 1, &done, &t);
        if (rc != VT_OK && first_rc == VT_OK) first_rc = rc;
        vt_request_result& res = results[i];
        res.admit_s = ((t.t_admit > 0 ? t.t_admit : t.t_done) - t0) * 1e-3;
        res.done_s = (t.t_done - t0) * 1e-3;
        res.latency_ms = res.done_s * 1e3 - reqs[i].arrival_s * 1e3;
        res.launches = t.launches;
    }
    int rc = worker_stop(sv, stats);
    if (first_rc != VT_OK) return first_rc;
    return rc;
}

}  // extern "C"
