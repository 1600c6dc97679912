// vt_serve.cpp -- render-request serving loop (no reference counterpart).
//
// Scenes are volumes whose every brick is resident in HBM, with a min/max
// pyramid over the octree levels so that any isovalue gets its membership
// tables without rebuilding the octree (a node is valid for iso t exactly
// when min < t <= max over its zero-padded samples, the closed form of the
// reference builder).  Requests arrive over time; the loop admits them while
// their output buffers fit the HBM budget and tile slots are free, and every
// launch renders 16x16 tiles of all admitted requests together.  The launch
// is a captured hipGraph of fixed shape, replayed once per batch with a new
// tile list.  Mode 1 renders one request per launch, in arrival order (the
// plain per-request baseline).
#include <math.h>

#include <map>
#include <thread>

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
    std::map<uint32_t, uint8_t*> tables;      // isovalue bits -> membership bytes
    std::map<uint32_t, int> root;
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

}  // namespace

struct vt_server {
    vt_context* ctx = nullptr;
    int64_t budget = 0;
    int max_tiles = 0;
    int max_jobs = 256;
    std::vector<std::unique_ptr<Scene>> scenes;
    Arena arena;
    DevBuf<ServeJob> d_jobs;
    DevBuf<ServeTile> d_work;
    ServeCounters* d_ctr = nullptr;
    PinBuf<ServeTile> h_work;
    PinBuf<ServeCounters> h_ctr;
    PinBuf<ServeJob> h_jobs;   // pinned staging, one entry per job slot (async copies read it later)
    // one captured launch per lane width 2^lg (lg = 0..kServeWidths-1)
    static constexpr int kServeWidths = 4;
    hipGraph_t graph[kServeWidths] = {};
    hipGraphExec_t exec[kServeWidths] = {};
    int64_t lane_budget = 1 << 21;   // lanes of a batch (rays x 2^lg); <= 0: fixed 2^-lane_budget lanes
    ~vt_server()
    {
        for (int k = 0; k < kServeWidths; ++k) {
            if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
        }
        for (auto& s : scenes) {
            if (s->d_pool) (void)hipFree(s->d_pool);
            if (s->d_slot) (void)hipFree(s->d_slot);
            if (s->d_submm) (void)hipFree(s->d_submm);
            for (auto& t : s->tables) (void)hipFree(t.second);
        }
        if (arena.base) (void)hipFree(arena.base);
        if (d_ctr) (void)hipFree(d_ctr);
        d_jobs.release();
        d_work.release();
        h_work.release();
        h_ctr.release();
        h_jobs.release();
    }
};

namespace {

int tree_levels3(int X, int Y, int Z)
{
    int d = std::max(X, std::max(Y, Z)), n = 0;
    while ((1 << n) < d) ++n;
    return n;
}

// membership bytes of levels 1..levelO for isovalue t (one byte per parent)
int scene_table(vt_server* sv, Scene& sc, float iso, uint8_t** out, int* root)
{
    uint32_t key;
    memcpy(&key, &iso, 4);
    auto it = sc.tables.find(key);
    if (it != sc.tables.end()) {
        *out = it->second;
        *root = sc.root[key];
        return VT_OK;
    }
    std::vector<uint8_t> bytes(std::max<size_t>(1, sc.table_bytes), 0);
    for (int l = 1; l <= sc.level_o; ++l) {
        const auto& mn = sc.mn[(size_t)l];
        const auto& mx = sc.mx[(size_t)l];
        for (size_t c = 0; c < mn.size(); ++c)
            if (mn[c] < iso && iso <= mx[c]) bytes[(size_t)(sc.level_off[(size_t)l] + (c >> 3))] |= (uint8_t)(1u << (c & 7));
    }
    int r = sc.mn[0][0] < iso && iso <= sc.mx[0][0];
    uint8_t* d = nullptr;
    HIP_TRY(hipMalloc((void**)&d, bytes.size()));
    HIP_TRY(hipMemcpyAsync(d, bytes.data(), bytes.size(), hipMemcpyHostToDevice, sv->ctx->stream));
    HIP_TRY(hipStreamSynchronize(sv->ctx->stream));
    sc.tables[key] = d;
    sc.root[key] = r;
    *out = d;
    *root = r;
    return VT_OK;
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
    if (!sv->arena.init((size_t)hbm_budget_bytes)) return fail(VT_E_NOMEM, "serving arena");
    if (sv->d_jobs.ensure((size_t)sv->max_jobs) != hipSuccess || sv->d_work.ensure((size_t)max_tiles) != hipSuccess ||
        sv->h_work.ensure((size_t)max_tiles) != hipSuccess || sv->h_ctr.ensure(1) != hipSuccess ||
        sv->h_jobs.ensure((size_t)sv->max_jobs) != hipSuccess)
        return fail(VT_E_NOMEM, "serving buffers");
    HIP_TRY(hipMalloc((void**)&sv->d_ctr, sizeof(ServeCounters)));
    HIP_TRY(hipMemset(sv->d_ctr, 0, sizeof(ServeCounters)));
    // the batch launch at each lane width, captured once and replayed: the
    // tile list and count change between replays, the shape never does
    for (int lg = 0; lg < vt_server::kServeWidths; ++lg) {
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        hipError_t le = launch_serve_pass(sv->d_jobs.p, sv->d_work.p, sv->d_ctr, max_tiles, lg, ctx->stream);
        hipError_t ce = hipStreamEndCapture(ctx->stream, &sv->graph[lg]);
        if (le != hipSuccess) return fail(VT_E_HIP, std::string("capture: ") + hipGetErrorString(le));
        if (ce != hipSuccess) return fail(VT_E_HIP, std::string("capture: ") + hipGetErrorString(ce));
        HIP_TRY(hipGraphInstantiate(&sv->exec[lg], sv->graph[lg], nullptr, nullptr, 0));
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
        (void)hipStreamSynchronize(sv->ctx->stream);
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

int vt_server_run(vt_server* sv, const vt_request* reqs, int64_t nreq, int32_t mode, vt_request_result* results,
                  vt_serve_stats* stats)
{
    if (!sv || (nreq && (!reqs || !results))) return fail(VT_E_ARG, "NULL argument");
    if (mode != 0 && mode != 1) return fail(VT_E_ARG, "mode must be 0 (batched) or 1 (one request per launch)");
    HIP_TRY(hipSetDevice(sv->ctx->device));
    hipStream_t s = sv->ctx->stream;
    // validate and order by arrival
    std::vector<int64_t> order((size_t)nreq);
    for (int64_t i = 0; i < nreq; ++i) {
        const vt_request& q = reqs[i];
        if (q.scene < 0 || q.scene >= (int32_t)sv->scenes.size()) return fail(VT_E_ARG, "unknown scene");
        const Scene& qs = *sv->scenes[(size_t)q.scene];
        if (!(q.iso >= qs.iso_lo && q.iso <= qs.iso_hi))
            return fail(VT_E_ARG, "isovalue outside the scene's range");
        if (q.cam.W <= 0 || q.cam.H <= 0 || (int64_t)q.cam.W * q.cam.H > (1 << 26)) return fail(VT_E_ARG, "bad screen");
        size_t need = (size_t)q.cam.W * q.cam.H * 48 + 512;
        if ((int64_t)need > sv->budget) return fail(VT_E_ARG, "a request's outputs exceed the HBM budget");
        order[(size_t)i] = i;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    struct Live {
        int64_t req;
        int job;
        size_t off, bytes;
        int tiles, next_tile;
        int launches;
    };
    std::vector<Live> live;
    std::vector<int> free_jobs;
    for (int j = sv->max_jobs - 1; j >= 0; --j) free_jobs.push_back(j);
    size_t next = 0;
    int64_t done = 0, launches = 0, tiles_total = 0;
    uint64_t rays = 0;
    double busy_ms = 0.0;
    const double t0 = now_ms();
    auto elapsed_s = [&]() { return (now_ms() - t0) * 1e-3; };
    while (done < nreq) {
        // admit arrived requests while budget, job slots (and, in mode 1, an idle device) allow
        while (next < order.size() && reqs[order[next]].arrival_s <= elapsed_s() && !free_jobs.empty() &&
               (mode == 0 || live.empty())) {
            const vt_request& q = reqs[order[next]];
            size_t bytes = (size_t)q.cam.W * q.cam.H * 48;
            size_t off = sv->arena.alloc(bytes);
            if (off == (size_t)-1) break;   // wait for outputs to drain
            Scene& sc = *sv->scenes[(size_t)q.scene];
            uint8_t* table = nullptr;
            int root = 0;
            if (int rc = scene_table(sv, sc, q.iso, &table, &root)) return rc;
            const int job = free_jobs.back();
            free_jobs.pop_back();
            ServeJob& J = sv->h_jobs.p[job];
            memcpy(&J.cam, &q.cam, sizeof J.cam);
            J.oct.child_bytes = table;
            for (int l = 0; l < 24; ++l) J.oct.level_off[l] = sc.level_off[(size_t)l];
            J.oct.n = sc.n;
            J.oct.final_level = sc.level_o;
            J.oct.root_valid = root;
            J.bv = BrickView{sc.d_slot, sc.d_pool, sc.S, sc.inc, sc.slot_elems, sc.level_c,
                             3 * (sc.level_o - sc.level_c), 1ull << (3 * sc.level_c), sc.d_submm, sc.G};
            J.sh = Shading{q.iso, q.light[0], q.light[1], q.light[2], q.color_scale};
            J.rgba = reinterpret_cast<float4*>(sv->arena.base + off);
            J.hits = reinterpret_cast<HitRecord*>(sv->arena.base + off + (size_t)q.cam.W * q.cam.H * 16);
            J.tiles_x = (q.cam.W + 15) / 16;
            HIP_TRY(hipMemcpyAsync(sv->d_jobs.p + job, &J, sizeof J, hipMemcpyHostToDevice, s));
            Live L;
            L.req = order[next];
            L.job = job;
            L.off = off;
            L.bytes = bytes;
            L.tiles = J.tiles_x * ((q.cam.H + 15) / 16);
            L.next_tile = 0;
            L.launches = 0;
            live.push_back(L);
            results[L.req].admit_s = elapsed_s();
            ++next;
        }
        if (live.empty()) {
            if (next < order.size()) {   // idle until the next arrival
                double wait = reqs[order[next]].arrival_s - elapsed_s();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(wait * 1e6)));
            }
            continue;
        }
        // fill the launch with tiles, oldest request first
        int nw = 0;
        for (auto& L : live) {
            while (L.next_tile < L.tiles && nw < sv->max_tiles) {
                sv->h_work.p[nw].job = (uint32_t)L.job;
                sv->h_work.p[nw].tile = (uint32_t)L.next_tile++;
                ++nw;
            }
            if (nw == sv->max_tiles) break;
        }
        sv->h_ctr.p->nwork = (uint32_t)nw;
        sv->h_ctr.p->err = 0;
        double l0 = now_ms();
        HIP_TRY(hipMemcpyAsync(sv->d_work.p, sv->h_work.p, (size_t)nw * sizeof(ServeTile), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(sv->d_ctr, sv->h_ctr.p, sizeof(ServeCounters), hipMemcpyHostToDevice, s));
        // lanes per ray: the widest that keeps the batch within the lane budget
        int lg = 0;
        if (sv->lane_budget > 0)
            while (lg + 1 < vt_server::kServeWidths && ((int64_t)nw * 256 << (lg + 1)) <= sv->lane_budget) ++lg;
        else
            lg = (int)std::min<int64_t>(-sv->lane_budget, vt_server::kServeWidths - 1);
        HIP_TRY(hipGraphLaunch(sv->exec[lg], s));
        HIP_TRY(hipMemcpyAsync(sv->h_ctr.p, sv->d_ctr, sizeof(ServeCounters), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        busy_ms += now_ms() - l0;
        if (sv->h_ctr.p->err) return fail(VT_E_RANGE, "serving pass: a sample fell outside its brick slot");
        ++launches;
        tiles_total += nw;
        // retire finished requests
        for (size_t k = 0; k < live.size();) {
            Live& L = live[k];
            L.launches++;
            if (L.next_tile < L.tiles) { ++k; continue; }
            const vt_request& q = reqs[L.req];
            size_t R = (size_t)q.cam.W * q.cam.H;
            if (q.h_rgba)
                HIP_TRY(hipMemcpyAsync(q.h_rgba, sv->arena.base + L.off, R * 16, hipMemcpyDeviceToHost, s));
            if (q.h_hits)
                HIP_TRY(hipMemcpyAsync(q.h_hits, sv->arena.base + L.off + R * 16, R * 32, hipMemcpyDeviceToHost, s));
            if (q.h_rgba || q.h_hits) HIP_TRY(hipStreamSynchronize(s));
            vt_request_result& res = results[L.req];
            res.done_s = elapsed_s();
            res.latency_ms = (res.done_s - q.arrival_s) * 1e3;
            res.launches = L.launches;
            rays += R;
            sv->arena.release(L.off, L.bytes);
            free_jobs.push_back(L.job);
            live.erase(live.begin() + (long)k);
            ++done;
        }
    }
    if (stats) {
        stats->requests = nreq;
        stats->launches = launches;
        stats->tiles = tiles_total;
        stats->rays = (int64_t)rays;
        stats->wall_s = elapsed_s();
        stats->busy_s = busy_ms * 1e-3;
    }
    return VT_OK;
}

}  // extern "C"
