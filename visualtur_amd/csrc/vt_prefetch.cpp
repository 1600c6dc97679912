// vt_prefetch.cpp -- the learned brick-visibility prefetcher (no reference
// counterpart): MLP parameters, training step, inference, and the prefetch
// of the highest-scoring bricks of an upcoming camera into the brick cache.
#include <math.h>

#include <random>

#include "vt_context.hpp"
#include "vt_mlp.hpp"

namespace {

constexpr int kIn = 12, kInPad = 32, kHid = 256;
// internal fp32 parameter layout (W1 padded to 32 input rows)
constexpr size_t kW1 = 0, kB1 = kW1 + (size_t)kInPad * kHid, kW2 = kB1 + kHid, kB2 = kW2 + (size_t)kHid * kHid,
                 kW3 = kB2 + kHid, kB3 = kW3 + (size_t)kHid * kHid, kW4 = kB3 + kHid, kB4 = kW4 + kHid,
                 kParams = kB4 + 1;
// exposed layout (W1 with its 12 real input rows)
constexpr size_t kExposed = kParams - (size_t)(kInPad - kIn) * kHid;

}  // namespace

struct vt_mlp {
    vt_context* ctx = nullptr;
    int64_t max_batch = 0;
    int step = 0;
    DevBuf<float> P, G, Mo, Vo;   // master params, grads, Adam moments
    DevBuf<bf16> Pb;              // BF16 copies (same layout)
    DevBuf<bf16> X, H1, H2, H3, dH1, dH2, dH3;
    DevBuf<float> z, dz, y, loss;
    DevBuf<float> feats;
    DevBuf<uint64_t> codes;
    // fused scoring: transposed BF16 hidden-layer weights (rebuilt after
    // every change of the parameters), candidate logits, and the filtered
    // results (absent candidates at or above the threshold, resident
    // candidates' logits by slot) with their pinned host copies
    DevBuf<bf16> WT;
    bool wt_stale = true;
    DevBuf<float> zc, want_z, slot_score;
    DevBuf<int32_t> want_idx;
    DevBuf<uint32_t> nwant;
    PinBuf<float> h_want_z, h_slot_score;
    PinBuf<int32_t> h_want_idx;
    PinBuf<uint32_t> h_nwant;
    // eviction staging of vt_prefetch (reused after the next call's first wait)
    PinBuf<int64_t> h_eidx;
    PinBuf<int32_t> h_eval;
    DevBuf<int64_t> d_eidx;
    DevBuf<int32_t> d_eval;
    ~vt_mlp()
    {
        P.release(); G.release(); Mo.release(); Vo.release(); Pb.release();
        X.release(); H1.release(); H2.release(); H3.release(); dH1.release(); dH2.release(); dH3.release();
        z.release(); dz.release(); y.release(); loss.release(); feats.release(); codes.release();
        WT.release(); zc.release(); want_z.release(); slot_score.release(); want_idx.release(); nwant.release();
        h_want_z.release(); h_slot_score.release(); h_want_idx.release(); h_nwant.release();
        h_eidx.release(); h_eval.release(); d_eidx.release(); d_eval.release();
    }
};

namespace {

int mlp_check(vt_mlp* m, int64_t N)
{
    if (!m) return fail(VT_E_ARG, "mlp is NULL");
    if (N <= 0 || N > m->max_batch) return fail(VT_E_ARG, "batch size must be in [1, max_batch]");
    return VT_OK;
}

// expose <-> internal parameter layouts
void to_internal(const float* in, std::vector<float>& out)
{
    out.assign(kParams, 0.0f);
    memcpy(out.data() + kW1, in, (size_t)kIn * kHid * 4);
    memcpy(out.data() + kB1, in + (size_t)kIn * kHid, (kParams - kB1) * 4);
}

void to_exposed(const std::vector<float>& in, float* out)
{
    memcpy(out, in.data() + kW1, (size_t)kIn * kHid * 4);
    memcpy(out + (size_t)kIn * kHid, in.data() + kB1, (kParams - kB1) * 4);
}

int refresh_bf16(vt_mlp* m)
{
    HIP_TRY(launch_to_bf16(m->P.p, m->Pb.p, kParams, m->ctx->stream));
    m->wt_stale = true;
    return VT_OK;
}

// the scoring kernel's transposed weights, from the current BF16 copies
int ensure_wt(vt_mlp* m)
{
    if (!m->wt_stale) return VT_OK;
    if (m->WT.ensure((size_t)kHid * kInPad + 2 * (size_t)kHid * kHid) != hipSuccess)
        return fail(VT_E_NOMEM, "scoring weights");
    HIP_TRY(launch_pack_wt(m->Pb.p + kW1, m->Pb.p + kW2, m->Pb.p + kW3, m->WT.p, m->ctx->stream));
    m->wt_stale = false;
    return VT_OK;
}

ScoreParams score_params()
{
    return ScoreParams{(int)kB1, (int)kB2, (int)kB3, (int)kW4, (int)kB4};
}

int forward(vt_mlp* m, const float* d_x, int N)
{
    hipStream_t s = m->ctx->stream;
    const bf16* Wb = m->Pb.p;
    const float* P = m->P.p;
    HIP_TRY(launch_pack_x(d_x, N, m->X.p, s));
    HIP_TRY(launch_gemm(false, false, EPI_BIAS_RELU_BF16, m->X.p, kInPad, Wb + kW1, kHid, nullptr, m->H1.p, kHid, N,
                        kHid, kInPad, 1, P + kB1, nullptr, 0, s));
    HIP_TRY(launch_gemm(false, false, EPI_BIAS_RELU_BF16, m->H1.p, kHid, Wb + kW2, kHid, nullptr, m->H2.p, kHid, N,
                        kHid, kHid, 1, P + kB2, nullptr, 0, s));
    HIP_TRY(launch_gemm(false, false, EPI_BIAS_RELU_BF16, m->H2.p, kHid, Wb + kW3, kHid, nullptr, m->H3.p, kHid, N,
                        kHid, kHid, 1, P + kB3, nullptr, 0, s));
    HIP_TRY(launch_logit(m->H3.p, P + kW4, P + kB4, N, kHid, m->z.p, s));
    return VT_OK;
}

// gradients of the mean weighted BCE into m->G (zeroed here); loss into m->loss[0]
int backward(vt_mlp* m, const float* d_y, int N, float pos_weight)
{
    hipStream_t s = m->ctx->stream;
    const bf16* Wb = m->Pb.p;
    const float* P = m->P.p;
    float* G = m->G.p;
    HIP_TRY(hipMemsetAsync(G, 0, kParams * 4, s));
    HIP_TRY(hipMemsetAsync(m->loss.p, 0, 4, s));
    HIP_TRY(launch_bce(m->z.p, d_y, N, pos_weight, m->dz.p, m->loss.p, s));
    HIP_TRY(launch_colsum(m->H3.p, m->dz.p, N, kHid, G + kW4, s));
    HIP_TRY(launch_sum(m->dz.p, N, G + kB4, s));
    HIP_TRY(launch_outer_mask(m->dz.p, P + kW4, m->H3.p, N, kHid, m->dH3.p, s));
    const int split = std::max(1, N / 512);
    // layer 3
    HIP_TRY(launch_gemm(true, false, EPI_ATOMIC_F32, m->H2.p, kHid, m->dH3.p, kHid, G + kW3, nullptr, kHid, kHid, kHid,
                        N, split, nullptr, nullptr, 0, s));
    HIP_TRY(launch_colsum(m->dH3.p, nullptr, N, kHid, G + kB3, s));
    HIP_TRY(launch_gemm(false, true, EPI_MASK_BF16, m->dH3.p, kHid, Wb + kW3, kHid, nullptr, m->dH2.p, kHid, N, kHid,
                        kHid, 1, nullptr, m->H2.p, kHid, s));
    // layer 2
    HIP_TRY(launch_gemm(true, false, EPI_ATOMIC_F32, m->H1.p, kHid, m->dH2.p, kHid, G + kW2, nullptr, kHid, kHid, kHid,
                        N, split, nullptr, nullptr, 0, s));
    HIP_TRY(launch_colsum(m->dH2.p, nullptr, N, kHid, G + kB2, s));
    HIP_TRY(launch_gemm(false, true, EPI_MASK_BF16, m->dH2.p, kHid, Wb + kW2, kHid, nullptr, m->dH1.p, kHid, N, kHid,
                        kHid, 1, nullptr, m->H1.p, kHid, s));
    // layer 1
    HIP_TRY(launch_gemm(true, false, EPI_ATOMIC_F32, m->X.p, kInPad, m->dH1.p, kHid, G + kW1, nullptr, kHid, kInPad,
                        kHid, N, split, nullptr, nullptr, 0, s));
    HIP_TRY(launch_colsum(m->dH1.p, nullptr, N, kHid, G + kB1, s));
    return VT_OK;
}

}  // namespace

extern "C" {

int vt_mlp_num_params(void) { return (int)kExposed; }

int vt_mlp_create(vt_context* ctx, uint64_t seed, int64_t max_batch, vt_mlp** out)
{
    if (!ctx || !out) return fail(VT_E_ARG, "NULL argument");
    if (max_batch <= 0 || max_batch > (1 << 24)) return fail(VT_E_ARG, "max_batch must be in [1, 2^24]");
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<vt_mlp> m(new vt_mlp);
    m->ctx = ctx;
    m->max_batch = max_batch;
    size_t N = (size_t)max_batch;
    bool ok = m->P.ensure(kParams) == hipSuccess && m->G.ensure(kParams) == hipSuccess &&
              m->Mo.ensure(kParams) == hipSuccess && m->Vo.ensure(kParams) == hipSuccess &&
              m->Pb.ensure(kParams) == hipSuccess && m->X.ensure(N * kInPad) == hipSuccess &&
              m->H1.ensure(N * kHid) == hipSuccess && m->H2.ensure(N * kHid) == hipSuccess &&
              m->H3.ensure(N * kHid) == hipSuccess && m->dH1.ensure(N * kHid) == hipSuccess &&
              m->dH2.ensure(N * kHid) == hipSuccess && m->dH3.ensure(N * kHid) == hipSuccess &&
              m->z.ensure(N) == hipSuccess && m->dz.ensure(N) == hipSuccess && m->y.ensure(N) == hipSuccess &&
              m->loss.ensure(1) == hipSuccess && m->feats.ensure(N * kIn) == hipSuccess &&
              m->codes.ensure(N) == hipSuccess;
    if (!ok) return fail(VT_E_NOMEM, "mlp buffers");
    // He-uniform initialisation of the weights, zero biases
    std::mt19937_64 rng(seed);
    std::vector<float> exposed(kExposed, 0.0f);
    auto fill = [&](size_t off, int fan_in, size_t count) {
        std::uniform_real_distribution<float> u(-sqrtf(6.0f / fan_in), sqrtf(6.0f / fan_in));
        for (size_t i = 0; i < count; ++i) exposed[off + i] = u(rng);
    };
    const size_t e_w1 = 0, e_w2 = (size_t)kIn * kHid + kHid, e_w3 = e_w2 + (size_t)kHid * kHid + kHid,
                 e_w4 = e_w3 + (size_t)kHid * kHid + kHid;
    fill(e_w1, kIn, (size_t)kIn * kHid);
    fill(e_w2, kHid, (size_t)kHid * kHid);
    fill(e_w3, kHid, (size_t)kHid * kHid);
    fill(e_w4, kHid, kHid);
    std::vector<float> internal;
    to_internal(exposed.data(), internal);
    HIP_TRY(hipMemcpy(m->P.p, internal.data(), kParams * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(m->Mo.p, 0, kParams * 4));
    HIP_TRY(hipMemset(m->Vo.p, 0, kParams * 4));
    if (int rc = refresh_bf16(m.get())) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *out = m.release();
    return VT_OK;
}

int vt_mlp_destroy(vt_mlp* m)
{
    if (m) {
        (void)hipStreamSynchronize(m->ctx->stream);
        delete m;
    }
    return VT_OK;
}

int vt_mlp_get_params(vt_mlp* m, float* out)
{
    if (!m || !out) return fail(VT_E_ARG, "NULL argument");
    std::vector<float> internal(kParams);
    HIP_TRY(hipMemcpyAsync(internal.data(), m->P.p, kParams * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    to_exposed(internal, out);
    return VT_OK;
}

int vt_mlp_set_params(vt_mlp* m, const float* in)
{
    if (!m || !in) return fail(VT_E_ARG, "NULL argument");
    std::vector<float> internal;
    to_internal(in, internal);
    HIP_TRY(hipMemcpyAsync(m->P.p, internal.data(), kParams * 4, hipMemcpyHostToDevice, m->ctx->stream));
    HIP_TRY(hipMemsetAsync(m->Mo.p, 0, kParams * 4, m->ctx->stream));
    HIP_TRY(hipMemsetAsync(m->Vo.p, 0, kParams * 4, m->ctx->stream));
    m->step = 0;
    if (int rc = refresh_bf16(m)) return rc;
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    return VT_OK;
}

int vt_mlp_forward(vt_mlp* m, const float* d_x, int64_t N, float* d_logits)
{
    if (int rc = mlp_check(m, N)) return rc;
    if (!d_x || !d_logits) return fail(VT_E_ARG, "NULL argument");
    if (int rc = forward(m, d_x, (int)N)) return rc;
    HIP_TRY(hipMemcpyAsync(d_logits, m->z.p, (size_t)N * 4, hipMemcpyDeviceToDevice, m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    return VT_OK;
}

int vt_mlp_gradients(vt_mlp* m, const float* d_x, const float* d_y, int64_t N, float pos_weight, float* grad_out,
                     float* loss_out)
{
    if (int rc = mlp_check(m, N)) return rc;
    if (!d_x || !d_y || !grad_out) return fail(VT_E_ARG, "NULL argument");
    if (int rc = forward(m, d_x, (int)N)) return rc;
    if (int rc = backward(m, d_y, (int)N, pos_weight)) return rc;
    std::vector<float> internal(kParams);
    float loss = 0.0f;
    HIP_TRY(hipMemcpyAsync(internal.data(), m->G.p, kParams * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    HIP_TRY(hipMemcpyAsync(&loss, m->loss.p, 4, hipMemcpyDeviceToHost, m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    to_exposed(internal, grad_out);
    if (loss_out) *loss_out = loss / (float)N;
    return VT_OK;
}

int vt_mlp_train_step(vt_mlp* m, const float* d_x, const float* d_y, int64_t N, float lr, float pos_weight,
                      float* loss_out)
{
    if (int rc = mlp_check(m, N)) return rc;
    if (!d_x || !d_y) return fail(VT_E_ARG, "NULL argument");
    if (int rc = forward(m, d_x, (int)N)) return rc;
    if (int rc = backward(m, d_y, (int)N, pos_weight)) return rc;
    m->step++;
    m->wt_stale = true;
    HIP_TRY(launch_adam(m->P.p, m->G.p, m->Mo.p, m->Vo.p, m->Pb.p, (int)kParams, lr, 0.9f, 0.999f, 1e-8f, m->step,
                        m->ctx->stream));
    if (loss_out) {
        float loss = 0.0f;
        HIP_TRY(hipMemcpyAsync(&loss, m->loss.p, 4, hipMemcpyDeviceToHost, m->ctx->stream));
        HIP_TRY(hipStreamSynchronize(m->ctx->stream));
        *loss_out = loss / (float)N;
    }
    return VT_OK;
}

int vt_mlp_features(vt_mlp* m, const vt_camera* cam, int32_t nLevels, int32_t levelC, const uint64_t* d_codes,
                    int64_t N, float* d_x)
{
    if (int rc = mlp_check(m, N)) return rc;
    if (!cam || !d_codes || !d_x) return fail(VT_E_ARG, "NULL argument");
    if (levelC < 0 || levelC > nLevels || nLevels > 20) return fail(VT_E_ARG, "bad levels");
    Camera c;
    memcpy(&c, cam, sizeof c);
    HIP_TRY(launch_features(c, nLevels, levelC, d_codes, (int)N, d_x, m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    return VT_OK;
}

}  // extern "C"

int vt_mlp_score(vt_mlp* m, const vt_camera* cam, int32_t nLevels, int32_t levelC, const uint64_t* d_codes,
                 int64_t N, float* d_logits)
{
    if (!m || !cam || !d_codes || !d_logits) return fail(VT_E_ARG, "NULL argument");
    if (N <= 0 || N > INT32_MAX) return fail(VT_E_ARG, "N must be in [1, 2^31)");
    if (levelC < 0 || levelC > nLevels || nLevels > 20) return fail(VT_E_ARG, "bad levels");
    Camera c;
    memcpy(&c, cam, sizeof c);
    if (int rc = ensure_wt(m)) return rc;
    HIP_TRY(launch_score(c, nLevels, levelC, d_codes, (int)N, m->WT.p, m->P.p, score_params(), d_logits, ScoreFilter{},
                         m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    return VT_OK;
}

// ====================================================== learned prefetch

int load_async(vt_context* ctx, const uint32_t* req, size_t n, std::vector<uint32_t>& left, int64_t* loaded,
               int64_t* misses, bool publish);

extern "C" int vt_prefetch(vt_context* ctx, vt_mlp* m, const vt_camera* cam, int64_t budget, float min_prob,
                           int64_t* loaded)
{
    return prefetch_bricks(ctx, m, cam, budget, min_prob, true, loaded);
}

namespace {
// With `publish` (between frames) the prefetcher may evict lower-scoring
// residents for its loads; without it (beside the render launches of frames
// in flight) it only fills free slots, since a resident brick may be one
// that a parked ray of a frame in flight is waiting for.
}  // namespace

int prefetch_bricks(vt_context* ctx, vt_mlp* m, const vt_camera* cam, int64_t budget, float min_prob, bool publish,
                    int64_t* loaded)
{
    if (!ctx || !m || !cam) return fail(VT_E_ARG, "NULL argument");
    if (!ctx->octree_ready) return fail(VT_E_STATE, "no octree uploaded");
    if (!ctx->cache) return fail(VT_E_STATE, "no cache");
    BrickCache& c = *ctx->cache;
    if (!c.d_vol && !c.host) return fail(VT_E_STATE, "no brick source attached");
    if (c.level_c > ctx->max_level) return fail(VT_E_STATE, "the octree must be uploaded down to levelC");
    if (budget < 0 || min_prob < 0.0f || min_prob >= 1.0f) return fail(VT_E_ARG, "bad budget or threshold");
    HIP_TRY(hipSetDevice(ctx->device));
    c.n = ctx->n;
    if (int rc = host_tier_start(ctx)) return rc;
    hipStream_t s = ctx->stream;
    // candidates: the bricks that hold octree nodes (the level-levelC run list)
    if (ctx->cand_level != c.level_c) {
        ctx->cand_codes.clear();
        const auto& R = ctx->up_runs[(size_t)c.level_c];
        for (size_t j = 0; j + 1 < R.size(); j += 2)
            for (uint64_t code = R[j]; code <= R[j + 1]; ++code) ctx->cand_codes.push_back(code);
        if (ctx->d_cand.ensure(std::max<size_t>(1, ctx->cand_codes.size())) != hipSuccess)
            return fail(VT_E_NOMEM, "candidate list");
        if (!ctx->cand_codes.empty())
            HIP_TRY(hipMemcpy(ctx->d_cand.p, ctx->cand_codes.data(), ctx->cand_codes.size() * 8,
                              hipMemcpyHostToDevice));
        ctx->cand_level = c.level_c;
    }
    const int64_t NC = (int64_t)ctx->cand_codes.size();
    int64_t nloaded = 0;
    if (NC > 0 && budget > 0) {
        // score every candidate in one fused launch; the kernel filters them
        // against the page table, which must first take in the page-table
        // writes of loads still publishing on the copy stream
        if (NC > INT32_MAX) return fail(VT_E_STATE, "too many candidate bricks");
        Camera k;
        memcpy(&k, cam, sizeof k);
        const size_t nc = (size_t)NC, ns = (size_t)c.gpu_slots;
        if (m->zc.ensure(nc) != hipSuccess || m->want_z.ensure(nc) != hipSuccess ||
            m->want_idx.ensure(nc) != hipSuccess || m->nwant.ensure(1) != hipSuccess ||
            m->slot_score.ensure(ns) != hipSuccess || m->h_want_z.ensure(nc) != hipSuccess ||
            m->h_want_idx.ensure(nc) != hipSuccess || m->h_nwant.ensure(1) != hipSuccess ||
            m->h_slot_score.ensure(ns) != hipSuccess)
            return fail(VT_E_NOMEM, "scoring buffers");
        if (int rc = ensure_wt(m)) return rc;
        const uint64_t base = 1ull << (3 * c.level_c);
        const float zmin = logf(min_prob / (1.0f - min_prob));   // logit threshold (-inf at 0)
        HIP_TRY(hipEventRecord(ctx->ev_copy, ctx->cstream));
        HIP_TRY(hipStreamWaitEvent(s, ctx->ev_copy, 0));
        HIP_TRY(hipMemsetAsync(m->nwant.p, 0, sizeof(uint32_t), s));
        HIP_TRY(launch_fill_f32(m->slot_score.p, (int)ns, -INFINITY, s));   // slots without a candidate
        ScoreFilter f;
        f.table = c.d_table;
        f.base = base;
        f.zmin = zmin;
        f.slot_score = m->slot_score.p;
        f.nwant = m->nwant.p;
        f.want_idx = m->want_idx.p;
        f.want_z = m->want_z.p;
        HIP_TRY(launch_score(k, ctx->n, c.level_c, ctx->d_cand.p, (int)NC, m->WT.p, m->P.p, score_params(), m->zc.p, f,
                             s));
        // the count, the head of the wanted list and the slot scores in one wait
        const size_t head = std::min<size_t>(nc, 16384);
        HIP_TRY(hipMemcpyAsync(m->h_nwant.p, m->nwant.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(m->h_want_idx.p, m->want_idx.p, head * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(m->h_want_z.p, m->want_z.p, head * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(m->h_slot_score.p, m->slot_score.p, ns * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        const size_t nw = std::min<size_t>(*m->h_nwant.p, nc);
        if (nw > head) {
            HIP_TRY(hipMemcpyAsync(m->h_want_idx.p + head, m->want_idx.p + head, (nw - head) * 4,
                                   hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(m->h_want_z.p + head, m->want_z.p + head, (nw - head) * 4, hipMemcpyDeviceToHost,
                                   s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        // absent candidates above the threshold (logit, candidate index), taken
        // best first (ties by candidate index, so the order does not depend on
        // the order in which the kernel appended them) from a heap, as far as
        // the loads reach: a few hundred of possibly thousands
        std::vector<std::pair<float, int32_t>> want;
        want.reserve(nw);
        for (size_t q = 0; q < nw; ++q) {
            const int32_t i = m->h_want_idx.p[q];
            if (c.table[(size_t)(ctx->cand_codes[(size_t)i] - base)] < 0) want.emplace_back(m->h_want_z.p[q], i);
        }
        const auto below = [](const std::pair<float, int32_t>& a, const std::pair<float, int32_t>& b) {
            return a.first < b.first || (a.first == b.first && a.second > b.second);
        };
        std::make_heap(want.begin(), want.end(), below);
        std::vector<std::pair<float, int32_t>> best;   // best first
        auto want_end = want.end();
        auto take = [&](size_t k) {   // extend `best` to k entries
            while (best.size() < k && want_end != want.begin()) {
                std::pop_heap(want.begin(), want_end, below);
                --want_end;
                best.push_back(*want_end);
            }
        };
        int64_t n = std::min<int64_t>((int64_t)want.size(), budget);
        if (!publish) n = std::min<int64_t>(n, c.available());   // frames in flight: free slots only
        if (n > c.available()) {
            // make room: evict resident bricks, lowest logit first (ties by
            // slot; bricks that are not candidates score -inf), each only for
            // a better candidate and only while its own probability is below
            // 1 - min_prob
            if (int rc = copy_stream_idle(ctx)) return rc;
            const float* slot_score = m->h_slot_score.p;
            std::vector<int64_t> res;
            res.reserve((size_t)c.used);
            for (int64_t sl = 0; sl < c.used; ++sl)
                if (c.slot_brick[(size_t)sl] >= 0) res.push_back(sl);
            const auto above = [&](int64_t x, int64_t y) {
                return slot_score[(size_t)x] > slot_score[(size_t)y] ||
                       (slot_score[(size_t)x] == slot_score[(size_t)y] && x > y);
            };
            std::make_heap(res.begin(), res.end(), above);
            if (m->h_eidx.ensure(ns) != hipSuccess || m->h_eval.ensure(ns) != hipSuccess ||
                m->d_eidx.ensure(ns) != hipSuccess || m->d_eval.ensure(ns) != hipSuccess)
                return fail(VT_E_NOMEM, "eviction staging");
            size_t ne = 0;
            auto res_end = res.end();
            while (c.available() < n && res_end != res.begin()) {
                const int64_t v = res.front();
                take((size_t)c.available() + 1);   // the candidate this slot would take
                const float zc = best[(size_t)c.available()].first;
                if (!(zc > slot_score[(size_t)v]) || !(slot_score[(size_t)v] < -zmin)) break;
                std::pop_heap(res.begin(), res_end, above);
                --res_end;
                const int64_t old = c.slot_brick[(size_t)v];
                c.table[(size_t)old] = -1;
                c.slot_brick[(size_t)v] = -1;
                c.ref[(size_t)v] = 0;
                c.free_list.push_back(v);
                c.stat_evictions++;
                m->h_eidx.p[ne] = old;
                m->h_eval.p[ne] = -1;
                ++ne;
            }
            if (ne > 0) {
                // the staging is pinned: the host rewrites it only after the
                // next call's first wait on this stream
                HIP_TRY(hipMemcpyAsync(m->d_eidx.p, m->h_eidx.p, ne * 8, hipMemcpyHostToDevice, s));
                HIP_TRY(hipMemcpyAsync(m->d_eval.p, m->h_eval.p, ne * 4, hipMemcpyHostToDevice, s));
                HIP_TRY(launch_scatter_i32(c.d_table, m->d_eidx.p, m->d_eval.p, (int)ne, s));
            }
            n = std::min<int64_t>(n, c.available());
        }
        take((size_t)n);
        // into free slots, on the copy stream, without waiting: the next
        // render pass runs while these bricks arrive
        std::vector<uint32_t> load((size_t)n), left;
        for (int64_t q = 0; q < n; ++q)
            load[(size_t)q] = (uint32_t)(ctx->cand_codes[(size_t)best[(size_t)q].second] - base);
        int64_t misses = 0;
        if (int rc = load_async(ctx, load.data(), load.size(), left, &nloaded, &misses, publish)) return rc;
    }
    if (loaded) *loaded = nloaded;
    return VT_OK;
}
