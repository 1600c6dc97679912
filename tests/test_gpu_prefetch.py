"""The learned prefetcher: trained on the GPU from the plain renderer's own
visited-brick traces, it must cut demand misses on an unseen camera path
without changing a single hit."""
import numpy as np
import pytest

from conftest import blobs_volume

pytestmark = pytest.mark.gpu


def test_prefetch_cuts_misses_and_keeps_hits(vt, vto, ctx):
    from visualtur_amd import prefetch as pf

    vol = blobs_volume(128, 48, seed=91)
    iso, level_o, level_c = 0.3, 4, 3
    n, runs = vto.build_octree(vol, iso, literal=False)
    flat = np.concatenate(runs[: level_o + 1]).astype(np.uint64)
    ctx.octree_upload(flat, [r.size for r in runs[: level_o + 1]], n, level_o)
    d_vol = ctx.to_device(vol)
    cand = pf.candidate_codes(runs[level_c])
    # traces are taken with every brick resident
    ctx.cache_create(level_c, 1 << (n - level_c), 3, int(cand.size), 0)
    ctx.cache_attach_device_volume(d_vol, vol.shape)
    prm = vt.render_params(level_o, level_c, iso)
    W = H = 128
    radius = (0.6, 1.0)   # close views: a frame sees under half of the bricks
    train_cams = [c for s in range(24) for c in pf.random_walk(W, H, 128, 8, seed=1000 + s, radius=radius)]
    traces = pf.collect_traces(ctx, train_cams, prm, level_c)
    net = vt.MLP(ctx, seed=7, max_batch=8192)
    X, Y = pf.build_dataset(ctx, net, train_cams, traces, cand, n, level_c, neg_per_pos=3, seed=1)
    losses, pos_w = pf.train(ctx, net, X, Y, steps=400, batch=4096, lr=1e-3)
    test_cams = pf.random_walk(W, H, 128, 48, seed=7, radius=radius)
    work = int(np.median([tr.size for tr in pf.collect_traces(ctx, test_cams, prm, level_c)]))
    # a device cache with room for a typical frame's working set and the
    # bricks about to enter the view, but not for the volume
    slots = min(int(cand.size) - 1, int(1.6 * work) + 8)
    print(f"candidates {cand.size} working set {work} slots {slots} samples {X.shape[0]} "
          f"loss {losses[:10].mean():.4f} -> {losses[-10:].mean():.4f}", flush=True)
    assert losses[-20:].mean() < 0.7 * losses[:20].mean()
    ctx.cache_create(level_c, 1 << (n - level_c), 3, slots, 0)
    ctx.cache_attach_device_volume(d_vol, vol.shape)
    ctx.set_early_loads(0)   # demand misses are compared: no loads that depend on timing
    try:
        plain, (rgba_p, hits_p) = pf.run_path(ctx, test_cams, prm)
        ctx.cache_clear()
        pref, (rgba_q, hits_q) = pf.run_path(ctx, test_cams, prm, net, budget=slots, min_prob=0.8)
    finally:
        ctx.set_early_loads()
    # identical pictures and hit records
    assert hits_p.tobytes() == hits_q.tobytes()
    assert np.array_equal(rgba_p.view(np.uint32), rgba_q.view(np.uint32))
    # and the same as the oracle
    rc, hc, _ = vto.render(vto.Volume(vol), vto.Octree(runs, n, level_o), test_cams[-1],
                           vto.params(level_o, level_c, iso))
    assert np.array_equal(hits_q["node"], hc["node"]) and np.array_equal(hits_q["voxel"], hc["voxel"])
    m_plain = sum(r["demand_misses"] for r in plain[1:])
    m_pref = sum(r["demand_misses"] for r in pref[1:])
    p_plain = np.mean([r["passes"] for r in plain[1:]])
    p_pref = np.mean([r["passes"] for r in pref[1:]])
    print(f"demand misses plain {m_plain} prefetch {m_pref}; passes {p_plain:.2f} -> {p_pref:.2f}; "
          f"prefetched {sum(r['prefetched'] for r in pref)}", flush=True)
    assert m_pref < 0.7 * m_plain
    net.close()
    d_vol.free()


def _threshold(z, q):
    """A logit threshold near quantile q with no logit (nor negated logit)
    within 1e-3 of it or of its negation, and its probability."""
    zs = np.sort(z.astype(np.float64))
    for qq in np.linspace(q, q + 0.2, 41):
        t = float(np.quantile(zs, qq))
        if np.min(np.abs(zs - t)) > 1e-3 and np.min(np.abs(zs + t)) > 1e-3:
            return t, float(1.0 / (1.0 + np.exp(-t)))
    raise AssertionError("no clean threshold")


def test_prefetch_selection_follows_the_scores(vt, vto, ctx):
    """vt_prefetch loads exactly the bricks its rule names, counted against
    the rule applied on the host to the logits of vt_mlp_score: the absent
    candidates at or above the threshold, best first, up to the budget; into
    free slots, then in place of residents (lowest logit first) while the
    resident scores below both the candidate it makes room for and
    logit(1 - min_prob)."""
    from visualtur_amd import prefetch as pf

    vol = blobs_volume(128, 48, seed=91)
    iso, level_o, level_c = 0.3, 4, 3
    n, runs = vto.build_octree(vol, iso, literal=False)
    flat = np.concatenate(runs[: level_o + 1]).astype(np.uint64)
    ctx.octree_upload(flat, [r.size for r in runs[: level_o + 1]], n, level_o)
    d_vol = ctx.to_device(vol)
    cand = pf.candidate_codes(runs[level_c])
    net = vt.MLP(ctx, seed=21, max_batch=4096)
    cam = vt.make_camera(96, 64, (150.0, 40.0, -60.0), *vt.look_at((150.0, 40.0, -60.0), (64, 64, 64)))
    d_c, d_z = ctx.to_device(cand), ctx.alloc(cand.size * 4)
    net.score(cam, n, level_c, d_c, cand.size, d_z)
    z = d_z.download(np.float32, (cand.size,)).astype(np.float64)
    assert np.unique(z).size == z.size
    order = np.lexsort((np.arange(z.size), -z))   # best first, ties by candidate index
    try:
        # free slots only: budget-limited, then the rest, then nothing
        t, p = _threshold(z, 0.6)
        above = int((z >= t).sum())
        assert 8 < above < cand.size
        ctx.cache_create(level_c, 1 << (n - level_c), 3, int(cand.size), 0)
        ctx.cache_attach_device_volume(d_vol, vol.shape)
        budget = above // 3
        assert net.prefetch(cam, budget, p) == budget
        assert net.prefetch(cam, cand.size, p) == above - budget
        assert net.prefetch(cam, cand.size, p) == 0
        assert ctx.cache_stats()["gpu_resident"] == above
        # a full cache holding the lowest-scoring candidates: loads replace
        # residents under the rule (a threshold at which it evicts some)
        slots = int(cand.size) // 3
        low = order[::-1][:slots]
        resident = set(low.tolist())

        def rule(t):
            want = [i for i in order if z[i] >= t and i not in resident]
            budget = len(want) - 3
            nw, avail = min(len(want), budget), 0
            for v in sorted(resident, key=lambda i: z[i]):
                if avail >= nw or not (z[want[avail]] > z[v] and z[v] < -t):
                    break
                avail += 1
            return want, budget, min(nw, avail), avail

        zs = np.sort(z)
        for q in np.linspace(0.2, 0.95, 76):
            t = float(np.quantile(zs, q))
            if np.min(np.abs(zs - t)) <= 1e-3 or np.min(np.abs(zs + t)) <= 1e-3:
                continue
            want, budget, expect, evicted = rule(t)
            if budget > 0 and evicted > 0:
                break
        else:
            raise AssertionError("no threshold at which the rule evicts")
        p = float(1.0 / (1.0 + np.exp(-t)))
        ctx.cache_create(level_c, 1 << (n - level_c), 3, slots, 0)
        ctx.cache_attach_device_volume(d_vol, vol.shape)
        ctx.cache_load_bricks(cand[low])
        ev0 = ctx.cache_stats()["evictions"]
        print(f"candidates {cand.size} slots {slots} wanted {len(want)} expect {expect} evictions {evicted}",
              flush=True)
        assert net.prefetch(cam, budget, p) == expect
        assert ctx.cache_stats()["evictions"] - ev0 == evicted
        assert ctx.cache_stats()["gpu_resident"] == slots
    finally:
        net.close()
        d_c.free()
        d_z.free()
        d_vol.free()
