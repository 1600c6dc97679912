// vt_frames.cpp -- frames in flight through the out-of-core brick cache
// (no reference counterpart; the reference renders one frame at a time,
// visualTur.cu:167-228, each pass waiting for the cache update of
// lruCache.cu:168-313 before the next).
//
// A worker thread serves frame requests as they are submitted (continuous
// batching over the context's octree and brick cache).  Up to max_frames
// frames are in flight: every launch carries the parked rays of the frames
// already started (their later passes) together with the first pass of the
// frames admitted since the last launch, so a new frame's first pass
// overlaps the tail of the frames before it, and the service between two
// launches (request list, evictions, loads) serves all of them at once.
// Rays keep their per-ray state at (frame slot << rshift | pixel); a frame
// is done when a launch that followed its first pass parks none of its rays.
// With a prefetcher set, the next waiting frame's bricks are scored and
// copied into free slots at each boundary, landing while the next launch
// runs, and published at the boundary after it.  Every frame's pixels and
// hit records are those vt_render_frame gives for the same camera.
#include <deque>
#include <unordered_map>

#include "vt_context.hpp"

namespace {

struct Pending {
    int64_t ticket;
    vt_camera cam;
    vt_render_params p;
    float* rgba;
    vt_hit* hits;
};

struct Ticket {
    int state = 0;   // 0 waiting, 1 in flight, 2 done
    int rc = VT_OK;
    std::string msg;
    vt_frame_stats st{};
    double t_submit = 0.0;
};

// A frame still parked after this many launches has stopped progressing
// (its bricks cannot be held): it fails instead of the worker spinning.
constexpr int kMaxLaunchesPerFrame = 100000;

struct Slot {
    bool used = false;
    bool fresh = false;     // first pass not launched yet
    int64_t ticket = 0;
    uint32_t err = 0;
    vt_frame_stats st{};
};

}  // namespace

struct vt_frames {
    vt_context* ctx = nullptr;
    int max_frames = 0;
    int rshift = 0;
    int level_o = -1, level_c = -1;
    int defer_cap = 0;                 // deferral entries per ray the state holds
    DevBuf<uint64_t> node, defer;
    DevBuf<uint32_t> list[2];
    DevBuf<uint8_t> phase;
    DevBuf<RenderJob> d_jobs;
    PinBuf<RenderJob> h_jobs;
    DevBuf<uint32_t> d_parked;
    PinBuf<uint32_t> h_parked;
    // prefetcher
    vt_mlp* mlp = nullptr;
    int64_t pf_budget = 0;
    float pf_prob = 0.8f;
    // online state
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Pending> queue;
    std::unordered_map<int64_t, Ticket> tickets;
    int64_t next_ticket = 1;
    std::thread worker;
    bool stopping = false;
    int worker_rc = VT_OK;
    std::string worker_err;
    // statistics since the last vt_frames_get_stats
    vt_frames_stats stats{};
    double st_t0 = 0.0, busy_t0 = 0.0;
    bool busy = false;

    ~vt_frames()
    {
        node.release(); defer.release(); list[0].release(); list[1].release(); phase.release();
        d_jobs.release(); h_jobs.release(); d_parked.release(); h_parked.release();
    }
};

namespace {

void finish(vt_frames* f, Slot& s, int rc)
{
    std::lock_guard<std::mutex> g(f->m);
    Ticket& t = f->tickets[s.ticket];
    t.state = 2;
    t.rc = rc;
    if (rc != VT_OK) t.msg = vt_rt::g_err;
    s.st.ms_total = now_ms() - t.t_submit;
    t.st = s.st;
    f->stats.frames++;
    f->cv_done.notify_all();
    s = Slot();
}

// One frame into slot `k`: its camera, shading and outputs go to the job
// table ahead of the launch that starts it.
int admit(vt_frames* f, const Pending& q, int k, Slot& s)
{
    RenderJob& J = f->h_jobs.p[k];
    memcpy(&J.cam, &q.cam, sizeof J.cam);
    J.sh = Shading{q.p.iso, q.p.light[0], q.p.light[1], q.p.light[2], q.p.color_scale};
    J.rgba = reinterpret_cast<float4*>(q.rgba);
    J.hits = reinterpret_cast<HitRecord*>(q.hits);
    // the pinned entry is rewritten only after its previous frame finished,
    // i.e. after every launch that read it
    HIP_TRY(hipMemcpyAsync(f->d_jobs.p + k, &J, sizeof J, hipMemcpyHostToDevice, f->ctx->stream));
    s = Slot();
    s.used = true;
    s.fresh = true;
    s.ticket = q.ticket;
    std::lock_guard<std::mutex> g(f->m);
    f->tickets[q.ticket].state = 1;
    return VT_OK;
}

// The worker's loop.  Returns a failure that ends it (every frame in flight
// or waiting then fails with it).
int frames_loop(vt_frames* f, std::deque<Pending>& waiting, std::vector<Slot>& slots)
{
    vt_context* ctx = f->ctx;
    if (int rc = set_device(ctx)) return rc;
    BrickCache& c = *ctx->cache;
    int cur = 0;
    uint32_t in_count = 0;   // parked rays of frames in flight, in list[cur]
    // per frame slot, its parked rays (its list at slot << rshift in list[cur])
    uint32_t in_counts[kMaxFrames] = {};
    bool first_launch = true;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(f->m);
            const bool any = std::any_of(slots.begin(), slots.end(), [](const Slot& s) { return s.used; });
            if (!any && waiting.empty() && in_count == 0)
                f->cv_work.wait(lk, [&] { return f->stopping || !f->queue.empty(); });
            while (!f->queue.empty()) {
                waiting.push_back(f->queue.front());
                f->queue.pop_front();
            }
            const bool any2 = std::any_of(slots.begin(), slots.end(), [](const Slot& s) { return s.used; });
            if (f->stopping && !any2 && waiting.empty()) return VT_OK;
            if (!f->busy && (any2 || !waiting.empty())) {
                f->busy = true;
                f->busy_t0 = now_ms();
            }
        }
        // admit waiting frames into free slots, in submission order
        for (int k = 0; k < f->max_frames && !waiting.empty(); ++k) {
            if (slots[(size_t)k].used) continue;
            if (int rc = admit(f, waiting.front(), k, slots[(size_t)k])) return rc;
            waiting.pop_front();
        }
        // the next waiting frame's bricks, landing while this launch runs:
        // into free slots only, which the clock provides (it never takes a
        // brick loaded at this boundary for the parked rays: those are
        // pinned in the current epoch)
        if (f->mlp && !waiting.empty() && f->pf_budget > 0) {
            int64_t loaded = 0;
            if (int rc = copy_stream_idle(ctx)) return rc;   // the copies of the last boundary have landed
            // at most a quarter of the cache for speculative loads
            const int64_t room = std::min<int64_t>({f->pf_budget, std::max<int64_t>(ctx->early_reserve, 1),
                                                    std::max<int64_t>(c.gpu_slots / 4, 1)});
            if (int rc = make_reserve(ctx, room)) return rc;
            if (int rc = prefetch_bricks(ctx, f->mlp, &waiting.front().cam, f->pf_budget, f->pf_prob, false, &loaded))
                return rc;
            f->stats.bricks_prefetched += loaded;
        }
        // the launch: parked rays of the frames in flight, then the first
        // pass of the fresh ones
        RenderArgs a;
        memset(&a, 0, sizeof a);
        int nfresh = 0;
        uint32_t tiles = 0;
        a.fresh_tile0[0] = 0;
        for (int k = 0; k < f->max_frames; ++k) {
            Slot& s = slots[(size_t)k];
            if (!s.used || !s.fresh) continue;
            const Camera& cam = f->h_jobs.p[k].cam;
            a.fresh_slot[nfresh] = (uint32_t)k;
            const uint32_t tx8 = (uint32_t)((cam.W + 7) / 8), ty8 = (uint32_t)((cam.H + 7) / 8);
            if (ctx->supertile > 1) {   // whole rounds of S x S blocks for the 8 XCDs (k_render_multi)
                const uint32_t S = (uint32_t)ctx->supertile, blocks = ((tx8 + S - 1) / S) * ((ty8 + S - 1) / S);
                tiles += (blocks + 7) / 8 * 8 * S * S;
            } else {
                tiles += tx8 * ty8;
            }
            a.fresh_tile0[++nfresh] = tiles;
        }
        if (nfresh == 0 && in_count == 0) {
            // nothing to launch: a frame in flight always has parked rays
            // or a first pass to run, so none is left here
            for (auto& s : slots)
                if (s.used) finish(f, s, s.err ? fail(VT_E_RANGE, "a sample fell outside its brick slot") : VT_OK);
            continue;
        }
        a.oct = make_view(ctx, f->level_o);
        a.level_c = c.level_c;
        a.shift = 3 * (f->level_o - f->level_c);
        a.brick_base = 1ull << (3 * c.level_c);
        a.brick_slot = c.d_table;
        a.pool = c.d_pool;
        a.S = c.S;
        a.inc = c.inc;
        a.slot_elems = c.slot_elems;
        a.ray_node = f->node.p;
        a.ray_phase = f->phase.p;
        a.ray_defer = f->defer_cap > 0 ? f->defer.p : nullptr;
        a.max_defer = std::min(ctx->max_defer, f->defer_cap);
        a.max_defer_late = std::min(ctx->max_defer_late, f->defer_cap);
        a.node_budget = ctx->first_nodes;
        a.req_bits = ctx->req_bits.p;
        a.req_list = ctx->req_list.p;
        a.slot_ref = c.d_ref;
        a.brick_trace = nullptr;
        a.submm = ctx->skip_empty ? c.d_submm : nullptr;
        a.G = c.G;
        // the proven sides follow the iso of the newest frame (frames in
        // flight with another iso read the (min, max))
        if (ctx->skip_empty && nfresh > 0) {
            if (int rc = ensure_sides(ctx, f->h_jobs.p[a.fresh_slot[nfresh - 1]].sh.iso)) return rc;
        }
        a.sides = ctx->skip_empty ? c.d_sides : nullptr;
        a.side_words = c.side_words;
        a.side_iso = c.side_iso;
        a.ctr = ctx->d_ctr;
        a.collect_stats = 0;
        a.jobs = f->d_jobs.p;
        a.rshift = f->rshift;
        a.defer_stride = (uint32_t)f->max_frames << f->rshift;
        a.job_parked = f->d_parked.p;
        a.nfresh = nfresh;
        a.supertile = ctx->supertile;
        a.in_list = in_count ? f->list[cur].p : nullptr;
        a.in_count = in_count;
        a.list_block0[0] = 0;
        for (int k = 0; k < kMaxFrames; ++k) {
            a.list_count[k] = k < f->max_frames ? in_counts[k] : 0u;
            a.list_block0[k + 1] = a.list_block0[k] + (a.list_count[k] + 63) / 64;
        }
        a.out_list = f->list[cur ^ 1].p;
        if (first_launch) {   // later launches follow a service that resets these
            HIP_TRY(hipMemsetAsync(ctx->d_ctr, 0, sizeof(FrameCounters), ctx->stream));
            HIP_TRY(hipMemsetAsync(ctx->req_bits.p, 0, ctx->req_bits.n * 4, ctx->stream));
            first_launch = false;
        }
        ctx->pass_state_clean = false;   // vt_render_frame resets the pass state after these frames
        HIP_TRY(hipMemsetAsync(f->d_parked.p, 0, (size_t)f->max_frames * 4, ctx->stream));
        // lanes per ray: a launch without a first pass spreads its few rays
        int lg = 0;
        if (nfresh == 0 && ctx->wide_lanes > 0)
            while (lg < 6 && ((int64_t)in_count << (lg + 1)) <= ctx->wide_lanes) ++lg;
        else if (nfresh == 0 && ctx->wide_lanes < 0)
            lg = (int)-ctx->wide_lanes;
        // free slots for the bricks a first pass requests while it runs
        const bool early = ctx->early_reserve > 0 && nfresh > 0;
        if (early) {
            if (int rc = make_reserve(ctx, std::min(ctx->early_reserve, ctx->early_last))) return rc;
        }
        // bricks loaded from here until the service after this launch (beside
        // it, and by that service) are held until the launch after it has run
        const uint64_t launch_epoch = c.epoch;
        HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
        HIP_TRY(launch_render_multi(a, lg, ctx->stream));
        HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
        HIP_TRY(hipMemcpyAsync(ctx->h_ctr, ctx->d_ctr, sizeof(FrameCounters), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(c.h_ref.p, c.d_ref, (size_t)c.gpu_slots, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(f->h_parked.p, f->d_parked.p, (size_t)f->max_frames * 4, hipMemcpyDeviceToHost,
                               ctx->stream));
        for (int k = 0; k < nfresh; ++k) slots[a.fresh_slot[k]].fresh = false;
        vt_frame_stats st;
        memset(&st, 0, sizeof st);
        if (early) {
            int64_t loaded = 0, misses = 0;
            if (int rc = load_during_pass(ctx, ctx->ev1, &loaded, &misses)) return rc;
            st.bricks_uploaded += loaded;
            st.bricks_early += loaded;
            st.host_tier_misses += misses;
        }
        // copies that landed beside the launch (early loads, prefetch) become
        // visible to the next one
        if (int rc = publish_loads(ctx, ctx->stream)) return rc;
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        f->stats.launches++;
        f->stats.ms_kernel += ms;
        uint32_t parked = 0;
        for (int k = 0; k < f->max_frames; ++k) {
            in_counts[k] = f->h_parked.p[k];
            parked += in_counts[k];
        }
        const uint32_t nreq = ctx->h_ctr->req_count;
        const uint32_t err = ctx->h_ctr->err;
        if (nreq > (uint32_t)c.nbricks) return fail(VT_E_STATE, "request list overflow");
        if (nfresh > 0) ctx->early_last = nreq;
        std::vector<uint32_t>& left = ctx->req_left;
        left.clear();
        uint32_t* rq = ctx->req_list.p;
        for (uint32_t i = 0; i < nreq; ++i) {
            const uint32_t v = __atomic_load_n(&rq[i], __ATOMIC_ACQUIRE);
            if (v == 0) return fail(VT_E_STATE, "request entry not visible after the pass");
            left.push_back(v - 1);
        }
        memset(rq, 0, (size_t)nreq * 4);
        // the service between launches: loads for every frame in flight; its
        // last kernel clears the request bitmap, counters and reference bytes.
        // The rays parked in this launch wait for bricks loaded beside it and
        // by this service: the clock holds both until the next launch ran.
        c.hold_from = launch_epoch;
        const PassReset reset{ctx->req_bits.p, (int64_t)ctx->req_bits.n, &ctx->d_ctr->req_count, 2, c.d_ref,
                              c.gpu_slots};
        if (!left.empty()) {
            if (int rc = service_requests(ctx, left.data(), (int64_t)left.size(), &st, &reset)) return rc;
        } else {
            if (int rc = fold_ref_bits(ctx, true)) return rc;
            HIP_TRY(launch_finish_service(c.d_table, nullptr, nullptr, 0, reset, ctx->stream));
        }
        if (err) HIP_TRY(hipMemsetAsync(&ctx->d_ctr->err, 0, 4, ctx->stream));
        f->stats.bricks_requested += nreq;
        f->stats.bricks_uploaded += st.bricks_uploaded;
        f->stats.bricks_early += st.bricks_early;
        f->stats.host_tier_misses += st.host_tier_misses;
        // frames whose started rays all finished are done
        for (int k = 0; k < f->max_frames; ++k) {
            Slot& s = slots[(size_t)k];
            if (!s.used || s.fresh) continue;
            if (s.st.passes >= kMaxLaunchesPerFrame)
                return fail(VT_E_STATE, "a frame in flight made no progress in " +
                                            std::to_string(kMaxLaunchesPerFrame) + " launches");
            s.st.passes++;
            s.st.ms_kernel += ms;
            s.st.rays_parked += f->h_parked.p[k];
            s.err |= err;
            if (f->h_parked.p[k] == 0) {
                c.epoch++;
                c.protect_from = c.epoch;
                finish(f, s, s.err ? fail(VT_E_RANGE, "a sample fell outside its brick slot") : VT_OK);
            }
        }
        cur ^= 1;
        in_count = parked;
        std::lock_guard<std::mutex> g(f->m);
        const bool any = std::any_of(slots.begin(), slots.end(), [](const Slot& x) { return x.used; });
        if (f->busy && !any && waiting.empty() && f->queue.empty()) {
            f->stats.busy_s += (now_ms() - f->busy_t0) * 1e-3;
            f->busy = false;
        }
    }
}

void frames_worker(vt_frames* f)
{
    std::deque<Pending> waiting;
    std::vector<Slot> slots((size_t)f->max_frames);
    const int rc = frames_loop(f, waiting, slots);
    if (f->ctx->cache) f->ctx->cache->hold_from = ~0ull;   // one frame at a time again
    if (rc != VT_OK) {
        const std::string msg = vt_rt::g_err;
        (void)hipStreamSynchronize(f->ctx->stream);
        (void)hipStreamSynchronize(f->ctx->cstream);
        std::lock_guard<std::mutex> g(f->m);
        f->worker_rc = rc;
        f->worker_err = msg;
        auto fail_ticket = [&](int64_t t) {
            Ticket& k = f->tickets[t];
            k.state = 2;
            k.rc = rc;
            k.msg = msg;
        };
        for (auto& s : slots)
            if (s.used) fail_ticket(s.ticket);
        for (auto& q : waiting) fail_ticket(q.ticket);
        for (auto& q : f->queue) fail_ticket(q.ticket);
        f->queue.clear();
        f->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int vt_frames_create(vt_context* ctx, int32_t max_frames, int64_t max_rays, vt_frames** out)
{
    if (!ctx || !out) return fail(VT_E_ARG, "NULL argument");
    if (max_frames < 1 || max_frames > kMaxFrames) return fail(VT_E_ARG, "max_frames must be in [1, 32]");
    if (max_rays < 1 || max_rays > (1ll << 26)) return fail(VT_E_ARG, "max_rays must be in [1, 2^26]");
    if (!ctx->octree_ready) return fail(VT_E_STATE, "no octree uploaded");
    if (!ctx->cache) return fail(VT_E_STATE, "no brick cache");
    if (!ctx->cache->d_vol && !ctx->cache->host) return fail(VT_E_STATE, "no brick source attached");
    if (ctx->persistent) return fail(VT_E_STATE, "the persistent pass option is not used with frames in flight");
    if (int rc = set_device(ctx)) return rc;
    std::unique_ptr<vt_frames> f(new vt_frames);
    f->ctx = ctx;
    f->max_frames = max_frames;
    while ((1ll << f->rshift) < max_rays) f->rshift++;
    if (((int64_t)max_frames << f->rshift) > (1ll << 31)) return fail(VT_E_ARG, "max_frames x max_rays too large");
    const size_t n = (size_t)max_frames << f->rshift;
    f->defer_cap = std::max(ctx->max_defer, ctx->max_defer_late);
    if (f->node.ensure(n) != hipSuccess || f->phase.ensure(n) != hipSuccess || f->list[0].ensure(n) != hipSuccess ||
        f->list[1].ensure(n) != hipSuccess || (f->defer_cap > 0 && f->defer.ensure(n * (size_t)f->defer_cap) != hipSuccess) ||
        f->d_jobs.ensure((size_t)max_frames) != hipSuccess || f->h_jobs.ensure((size_t)max_frames) != hipSuccess ||
        f->d_parked.ensure((size_t)max_frames) != hipSuccess || f->h_parked.ensure((size_t)max_frames) != hipSuccess)
        return fail(VT_E_NOMEM, "per-ray state of the frames in flight");
    // slot entries read before a frame is admitted there (idle lanes) hold zeros
    HIP_TRY(hipMemset(f->d_jobs.p, 0, (size_t)max_frames * sizeof(RenderJob)));
    BrickCache& c = *ctx->cache;
    c.n = ctx->n;
    if (int rc = host_tier_start(ctx)) return rc;
    f->st_t0 = now_ms();
    vt_frames* raw = f.release();
    raw->worker = std::thread(frames_worker, raw);
    *out = raw;
    return VT_OK;
}

int vt_frames_set_prefetch(vt_frames* f, vt_mlp* m, int64_t budget, float min_prob)
{
    if (!f) return fail(VT_E_ARG, "NULL argument");
    if (m && (budget < 0 || min_prob < 0.0f || min_prob >= 1.0f)) return fail(VT_E_ARG, "bad budget or threshold");
    std::lock_guard<std::mutex> g(f->m);
    const bool idle = f->queue.empty() && std::all_of(f->tickets.begin(), f->tickets.end(),
                                                      [](const auto& t) { return t.second.state == 2; });
    if (!idle) return fail(VT_E_STATE, "set the prefetcher while no frame is in flight");
    f->mlp = m;
    f->pf_budget = m ? budget : 0;
    f->pf_prob = min_prob;
    return VT_OK;
}

int vt_frames_submit(vt_frames* f, const vt_camera* cam, const vt_render_params* p, float* d_rgba, vt_hit* d_hits,
                     int64_t* ticket)
{
    if (!f || !p || !d_rgba || !d_hits || !ticket) return fail(VT_E_ARG, "NULL argument (d_hits is required)");
    if (int rc = check_camera(cam)) return rc;
    vt_context* ctx = f->ctx;
    if ((int64_t)cam->W * cam->H > (1ll << f->rshift)) return fail(VT_E_ARG, "frame larger than max_rays");
    if (p->levelO < 0 || p->levelO > ctx->max_level) return fail(VT_E_ARG, "levelO beyond the uploaded octree levels");
    if (p->levelC != ctx->cache->level_c || p->levelC > p->levelO)
        return fail(VT_E_ARG, "levelC must match the cache and be <= levelO");
    if ((1 << (ctx->n - ctx->cache->level_c)) != ctx->cache->cube)
        return fail(VT_E_ARG, "cubeDim must equal 2^(nLevels-levelC)");
    std::lock_guard<std::mutex> g(f->m);
    if (f->worker_rc != VT_OK) return fail(f->worker_rc, "the frame worker failed: " + f->worker_err);
    if (f->stopping) return fail(VT_E_STATE, "shutting down");
    if (f->level_o < 0) {
        f->level_o = p->levelO;
        f->level_c = p->levelC;
    } else if (p->levelO != f->level_o || p->levelC != f->level_c) {
        return fail(VT_E_ARG, "every frame in flight uses the same levelO and levelC");
    }
    const int64_t t = f->next_ticket++;
    Ticket& k = f->tickets[t];
    k.t_submit = now_ms();
    f->queue.push_back(Pending{t, *cam, *p, d_rgba, d_hits});
    f->cv_work.notify_one();
    *ticket = t;
    return VT_OK;
}

int vt_frames_poll(vt_frames* f, int64_t ticket, int32_t wait, int32_t* done, vt_frame_stats* stats)
{
    if (!f || !done) return fail(VT_E_ARG, "NULL argument");
    std::unique_lock<std::mutex> lk(f->m);
    auto it = f->tickets.find(ticket);
    if (it == f->tickets.end()) return fail(VT_E_ARG, "unknown ticket");
    if (wait) f->cv_done.wait(lk, [&] { return f->tickets[ticket].state == 2; });
    it = f->tickets.find(ticket);
    if (it->second.state != 2) {
        *done = 0;
        return VT_OK;
    }
    *done = 1;
    const Ticket t = it->second;
    f->tickets.erase(it);
    if (stats) *stats = t.st;
    return t.rc == VT_OK ? VT_OK : fail(t.rc, t.msg);
}

int vt_frames_get_stats(vt_frames* f, vt_frames_stats* out)
{
    if (!f || !out) return fail(VT_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> g(f->m);
    const double now = now_ms();
    *out = f->stats;
    out->wall_s = (now - f->st_t0) * 1e-3;
    if (f->busy) {
        out->busy_s += (now - f->busy_t0) * 1e-3;
        f->busy_t0 = now;
    }
    f->stats = vt_frames_stats{};
    f->st_t0 = now;
    return VT_OK;
}

int vt_frames_destroy(vt_frames* f)
{
    if (!f) return VT_OK;
    {
        std::lock_guard<std::mutex> g(f->m);
        f->stopping = true;
        f->cv_work.notify_all();
    }
    if (f->worker.joinable()) f->worker.join();
    (void)hipStreamSynchronize(f->ctx->stream);
    (void)hipStreamSynchronize(f->ctx->cstream);
    delete f;
    return VT_OK;
}

}  // extern "C"
