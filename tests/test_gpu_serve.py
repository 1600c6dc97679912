"""The serving loop: every request's result is the same alone, in a batched
mix, and as the oracle renders it; per-request isovalues come from the
min/max pyramid and equal the builder's octree for that isovalue."""
import numpy as np
import pytest

from conftest import blobs_volume

pytestmark = pytest.mark.gpu


def test_mixed_requests_same_alone_and_in_the_mix(vt, vto, ctx):
    from visualtur_amd.prefetch import random_walk

    vol_a = blobs_volume(64, 16, seed=5)
    vol_b = blobs_volume(64, 10, seed=6, dims=(64, 48, 40)) - 0.2
    sv = vt.Server(ctx, hbm_budget=64 << 20, max_tiles=96)
    da, db = ctx.to_device(vol_a), ctx.to_device(vol_b)
    sa = sv.add_scene(da, vol_a.shape, 4, 2)
    sb = sv.add_scene(db, vol_b.shape, 3, 3)
    reqs = []
    sizes = [(64, 64), (96, 80), (37, 23), (160, 120), (17, 200)]
    rng = np.random.default_rng(3)
    for k in range(14):
        W, H = sizes[k % len(sizes)]
        scene = sa if k % 3 else sb
        iso = [0.2, 0.45, 0.3][k % 3] if scene == sa else [0.0, 0.1][k % 2]
        cam = random_walk(W, H, 64, 1, seed=100 + k)[0]
        reqs.append({"scene": scene, "cam": cam, "iso": iso, "arrival_s": 0.001 * (k // 5), "keep": True})
    mix, st_mix, out_mix = sv.run(reqs, mode=0)
    alone, st_alone, out_alone = sv.run(reqs, mode=1)
    assert st_mix["launches"] < st_alone["launches"]   # requests shared launches
    assert max(r["launches"] for r in mix) >= 1
    for k, q in enumerate(reqs):
        (rg, hg), (ra, ha) = out_mix[k], out_alone[k]
        assert hg.tobytes() == ha.tobytes(), k   # bytewise: NaN depths (0/0 secant, Q12) match too
        assert np.array_equal(rg.view(np.uint32), ra.view(np.uint32)), k
        # the oracle, with the reference builder's octree for this isovalue
        vol = vol_a if q["scene"] == sa else vol_b
        lo, lc = (4, 2) if q["scene"] == sa else (3, 3)
        n, runs = vto.build_octree(vol, q["iso"], literal=False)
        rc, hc, _ = vto.render(vto.Volume(vol), vto.Octree(runs, n, lo), q["cam"], vto.params(lo, lc, q["iso"]))
        assert np.array_equal(hg["node"], hc["node"]) and np.array_equal(hg["voxel"], hc["voxel"]), k
        nan = np.isnan(rc)
        assert np.array_equal(np.isnan(rg), nan)
        assert np.all(np.abs(rg[~nan] - rc[~nan]) <= 1.0 / 255), k
    sv.close()
    da.free()
    db.free()


def test_lane_widths_give_the_same_bytes(vt, vto, ctx):
    """The serving launch with 1, 2, 4 and 8 lanes per ray (and the lane
    budget's choice) renders every request to the same bytes."""
    from visualtur_amd.prefetch import random_walk

    vol = blobs_volume(64, 16, seed=15)
    sv = vt.Server(ctx, hbm_budget=64 << 20, max_tiles=64)
    dv = ctx.to_device(vol)
    sc = sv.add_scene(dv, vol.shape, 4, 2)
    sizes = [(64, 64), (96, 80), (37, 23), (17, 200)]
    reqs = [{"scene": sc, "cam": random_walk(*sizes[k % 4], 64, 1, seed=300 + k)[0], "iso": [0.2, 0.35][k % 2],
             "arrival_s": 0.0, "keep": True} for k in range(8)]
    outs = {}
    try:
        for lanes in (0, -1, -2, -3, 1 << 21):
            sv.set_lanes(lanes)
            _, _, out = sv.run(reqs, mode=0)
            outs[lanes] = [(out[k][1].tobytes(), out[k][0].view(np.uint32).tobytes()) for k in range(len(reqs))]
    finally:
        sv.set_lanes()
    for lanes, rows in outs.items():
        assert rows == outs[0], lanes
    n, runs = vto.build_octree(vol, reqs[0]["iso"], literal=False)
    rc, hc, _ = vto.render(vto.Volume(vol), vto.Octree(runs, n, 4), reqs[0]["cam"], vto.params(4, 2, reqs[0]["iso"]))
    hg = np.frombuffer(outs[-3][0][0], vt.HIT_DTYPE)
    assert np.array_equal(hg["node"], hc["node"]) and np.array_equal(hg["voxel"], hc["voxel"])
    sv.close()
    dv.free()


def test_admission_respects_the_budget(vt, ctx):
    vol = blobs_volume(32, 6, seed=1)
    sv = vt.Server(ctx, hbm_budget=2 << 20, max_tiles=64)
    d = ctx.to_device(vol)
    s = sv.add_scene(d, vol.shape, 3, 2)
    from visualtur_amd.prefetch import random_walk
    # 128x128 needs 786 KB of outputs: at most two fit a 2 MB budget at once
    reqs = [{"scene": s, "cam": random_walk(128, 128, 32, 1, seed=k)[0], "iso": 0.3, "arrival_s": 0.0}
            for k in range(6)]
    res, st, _ = sv.run(reqs, mode=0)
    spans = sorted((r["admit_s"], r["done_s"]) for r in res)
    for t in np.linspace(0, max(e for _, e in spans), 200):
        assert sum(1 for a, e in spans if a <= t < e) <= 2
    with pytest.raises(vt.VTError):
        sv.run([{"scene": s, "cam": random_walk(1024, 1024, 32, 1, seed=0)[0], "iso": 0.3}], mode=0)
    sv.close()
    d.free()


def test_isovalue_range_scene(vt, vto, ctx):
    """A scene for isovalues in [0.25, 0.35] holds only the bricks whose
    min/max straddles the range, renders every request in the range to the
    same bytes as a scene holding every brick (and as the oracle), and
    refuses isovalues outside the range."""
    from visualtur_amd.prefetch import random_walk

    vol = blobs_volume(64, 16, seed=5)
    sv = vt.Server(ctx, hbm_budget=64 << 20, max_tiles=96)
    d = ctx.to_device(vol)
    full = sv.add_scene(d, vol.shape, 4, 3)
    part, resident = sv.add_scene_range(d, vol.shape, 4, 3, 0.25, 0.35)
    # the bricks whose zero-padded samples straddle the range, from the volume
    pad = np.zeros((65, 65, 65), np.float32)
    pad[:64, :64, :64] = vol
    want = 0
    for bx in range(8):
        for by in range(8):
            for bz in range(8):
                box = pad[8 * bx:8 * bx + 9, 8 * by:8 * by + 9, 8 * bz:8 * bz + 9]
                want += int(box.min() < 0.35 and box.max() >= 0.25)
    print(f"resident bricks {resident} of 512 (expected {want})", flush=True)
    assert resident == want and 0 < resident < 512
    reqs = []
    for k, iso in enumerate([0.25, 0.3, 0.35, 0.3]):
        cam = random_walk(96, 80, 64, 1, seed=300 + k)[0]
        for scene in (full, part):
            reqs.append({"scene": scene, "cam": cam, "iso": iso, "arrival_s": 0.0, "keep": True})
    _, _, out = sv.run(reqs, mode=0)
    for k in range(0, len(reqs), 2):
        (rf, hf), (rp, hp) = out[k], out[k + 1]
        assert hf.tobytes() == hp.tobytes() and np.array_equal(rf.view(np.uint32), rp.view(np.uint32)), k
    q = reqs[-1]
    n, runs = vto.build_octree(vol, q["iso"], literal=False)
    rc, hc, _ = vto.render(vto.Volume(vol), vto.Octree(runs, n, 4), q["cam"], vto.params(4, 3, q["iso"]))
    assert np.array_equal(out[len(reqs) - 1][1]["node"], hc["node"])
    assert (hc["node"] != 0).mean() > 0.05
    with pytest.raises(vt.VTError):
        sv.run([{"scene": part, "cam": reqs[0]["cam"], "iso": 0.5, "arrival_s": 0.0}], mode=0)
    sv.close()
    d.free()


def test_online_submissions_from_another_thread(vt, vto, ctx):
    """Requests submitted from a second host thread while the worker renders
    earlier ones are admitted into the running stream of batches (continuous
    batching: they are admitted before the earlier requests finish) and give
    the same bytes as when each is rendered alone."""
    import threading
    import time

    from visualtur_amd.prefetch import random_walk

    vol = blobs_volume(64, 16, seed=25)
    sv = vt.Server(ctx, hbm_budget=512 << 20, max_tiles=256)
    d = ctx.to_device(vol)
    sc = sv.add_scene(d, vol.shape, 4, 2)
    heavy = [{"scene": sc, "cam": random_walk(1024, 1024, 64, 1, seed=400 + k)[0], "iso": 0.25, "keep": True}
             for k in range(6)]
    light = [{"scene": sc, "cam": random_walk(*[(64, 64), (96, 80), (37, 23)][k % 3], 64, 1, seed=500 + k)[0],
              "iso": [0.2, 0.3][k % 2], "keep": True} for k in range(9)]
    alone, st_alone, out_alone = sv.run(heavy + light, mode=1)
    sv.start(0)
    t_sub, tickets = {}, {}
    try:
        for k, q in enumerate(heavy):
            t_sub[k] = time.perf_counter()
            tickets[k] = sv.submit(q)

        def other():
            for k, q in enumerate(light):
                t_sub[len(heavy) + k] = time.perf_counter()
                tickets[len(heavy) + k] = sv.submit(q)

        th = threading.Thread(target=other)
        th.start()
        th.join()
        res, outs = {}, {}
        for k, t in tickets.items():
            done, r, o = sv.poll(t, wait=True)
            assert done
            res[k], outs[k] = r, o
    finally:
        st = sv.stop()
    first_heavy_done = max(t_sub[k] + res[k]["done_s"] for k in range(len(heavy)))
    light_admitted = [t_sub[k] + res[k]["admit_s"] for k in range(len(heavy), len(heavy) + len(light))]
    print(f"online: {st['launches']} launches for {st['requests']} requests, light requests admitted "
          f"{(first_heavy_done - min(light_admitted)) * 1e3:.2f} ms before the heavy ones finished", flush=True)
    assert st["requests"] == len(tickets) and st["launches"] < st_alone["launches"]
    assert min(light_admitted) < first_heavy_done
    for k in range(len(heavy) + len(light)):
        (rg, hg), (ra, ha) = outs[k], out_alone[k]
        assert hg.tobytes() == ha.tobytes() and np.array_equal(rg.view(np.uint32), ra.view(np.uint32)), k
    q = light[0]
    n, runs = vto.build_octree(vol, q["iso"], literal=False)
    rc, hc, _ = vto.render(vto.Volume(vol), vto.Octree(runs, n, 4), q["cam"], vto.params(4, 2, q["iso"]))
    assert np.array_equal(outs[len(heavy)][1]["node"], hc["node"])
    sv.close()
    d.free()


def test_failed_request_leaves_the_server_usable(vt, ctx):
    """A scene without a halo makes the march read outside its slots: its
    requests fail with VT_E_RANGE, their outputs are released, and a request
    needing nearly the whole HBM budget is then served on the same server."""
    from visualtur_amd.prefetch import random_walk

    vol = blobs_volume(64, 16, seed=35)
    budget = 2 << 20
    sv = vt.Server(ctx, hbm_budget=budget, max_tiles=64)
    d = ctx.to_device(vol)
    bad = sv.add_scene(d, vol.shape, 4, 2, cube_inc=0)
    good = sv.add_scene(d, vol.shape, 4, 2)
    reqs = [{"scene": bad, "cam": random_walk(96, 96, 64, 1, seed=600 + k)[0], "iso": 0.25} for k in range(4)]
    with pytest.raises(vt.VTError) as e:
        sv.run(reqs, mode=0)
    assert e.value.code == -7   # VT_E_RANGE
    side = int(np.sqrt((budget - 4096) / 48)) - 1   # outputs of nearly the whole budget
    big = {"scene": good, "cam": random_walk(side, side, 64, 1, seed=7)[0], "iso": 0.25, "keep": True}
    res, st, out = sv.run([big], mode=0)
    assert st["rays"] == side * side and (out[0][1]["node"] != 0).any()
    sv.close()
    d.free()


def test_many_isovalues_stay_within_the_budget(vt, vto, ctx):
    """Each new isovalue's membership tables come from the HBM arena: with
    more distinct isovalues than the budget can hold tables for, unused tables
    are evicted (least recently used first), every request is served, and
    the tables held never exceed what the budget leaves."""
    from visualtur_amd.prefetch import random_walk

    vol = blobs_volume(128, 20, seed=45)
    table_bytes = sum(8 ** (l - 1) for l in range(1, 7))   # levels 1..6, one byte per parent
    budget = 1 << 20
    sv = vt.Server(ctx, hbm_budget=budget, max_tiles=64)
    d = ctx.to_device(vol)
    sc = sv.add_scene(d, vol.shape, 6, 4)
    isos = [0.05 + 0.01 * k for k in range(48)]
    reqs = [{"scene": sc, "cam": random_walk(48, 40, 128, 1, seed=700 + k)[0], "iso": iso, "arrival_s": 0.0,
             "keep": k % 12 == 0} for k, iso in enumerate(isos)]
    res, st, out = sv.run(reqs, mode=0)
    held_max = budget // ((table_bytes + 255) // 256 * 256)
    print(f"tables built {st['tables_built']}, held {st['tables_held']} (at most {held_max})", flush=True)
    assert st["requests"] == len(reqs) and st["rays"] == 48 * 40 * len(reqs)
    assert st["tables_built"] >= len(isos) and st["tables_held"] < len(isos) and st["tables_held"] <= held_max
    for k in out:
        q = reqs[k]
        n, runs = vto.build_octree(vol, q["iso"], literal=False)
        rc, hc, _ = vto.render(vto.Volume(vol), vto.Octree(runs, n, 6), q["cam"], vto.params(6, 4, q["iso"]))
        assert np.array_equal(out[k][1]["node"], hc["node"]) and np.array_equal(out[k][1]["voxel"], hc["voxel"]), k
    sv.close()
    d.free()
