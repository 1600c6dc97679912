"""Frames in flight through the out-of-core brick cache (vt_frames): every
frame of a camera path, rendered with several frames sharing launches and
the service between them, with and without the prefetcher, has the bytes
vt_render_frame gives for the same camera (and the oracle's hits)."""
import numpy as np
import pytest

from gpu_scenes import Scene, assert_bits, blobs_volume, orbit_camera

pytestmark = pytest.mark.gpu


def _frames(vt, ctx, sc, cams, max_frames, mlp=None, budget=0, min_prob=0.5):
    bufs = [(ctx.alloc(c.W * c.H * 16), ctx.alloc(c.W * c.H * 32)) for c in cams]
    out, stats = [], []
    with vt.Frames(ctx, max_frames=max_frames, max_rays=max(c.W * c.H for c in cams)) as fr:
        if mlp is not None:
            fr.set_prefetch(mlp, budget, min_prob)
        tickets = [fr.submit(c, sc.prm_g, *bufs[k]) for k, c in enumerate(cams)]
        for t in tickets:
            done, st = fr.poll(t)
            assert done
            stats.append(st)
        agg = fr.stats()
    for (r, h), c in zip(bufs, cams):
        R = c.W * c.H
        out.append((r.download(np.float32, (R, 4)), h.download(vt.HIT_DTYPE, (R,))))
        r.free()
        h.free()
    return out, stats, agg


def test_frames_in_flight_equal_single_frames(vt, vto, ctx):
    from visualtur_amd.prefetch import random_walk

    vol = blobs_volume(64, 14, seed=71)
    sc = Scene(vto, vol, 0.3, 3, 2)
    sc.setup_gpu(vt, ctx, slots=7)   # 64 bricks through 7 slots: frames wait for loads
    sizes = [(96, 80), (64, 64), (120, 50), (33, 47), (96, 80), (80, 96), (64, 64), (17, 129)]
    cams = [random_walk(W, H, 64, 1, seed=900 + k, radius=(1.4, 2.2))[0] for k, (W, H) in enumerate(sizes)]
    single, passes = [], 0
    for c in cams:
        rg, hg, st = sc.render_gpu(vt, ctx, c)
        single.append((rg, hg))
        passes += st.passes
    net = vt.MLP(ctx, seed=3, max_batch=4096)
    try:
        for max_frames, pre in ((1, False), (2, False), (4, False), (8, False), (16, False), (32, False), (3, True)):
            ctx.cache_clear()
            # min_prob 0 with the untrained network: every absent candidate qualifies, so
            # the prefetcher loads at every boundary where a frame waits
            out, stats, agg = _frames(vt, ctx, sc, cams, max_frames, net if pre else None, budget=4, min_prob=0.0)
            print(f"frames in flight {max_frames} prefetch {pre}: {agg['launches']} launches for {len(cams)} frames "
                  f"({passes} passes one frame at a time), {agg['bricks_uploaded']} uploads, "
                  f"{agg['bricks_prefetched']} prefetched", flush=True)
            for k in range(len(cams)):
                assert out[k][1].tobytes() == single[k][1].tobytes(), (max_frames, k)
                assert out[k][0].tobytes() == single[k][0].tobytes(), (max_frames, k)
            assert agg["frames"] == len(cams) and all(s.passes >= 1 for s in stats)
            if max_frames > 1:
                assert agg["launches"] < passes
            if pre:
                assert agg["bricks_prefetched"] > 0
    finally:
        net.close()
        ctx.cache_destroy()
    rc, hc = sc.render_cpu(vto, cams[0])
    assert_bits(single[0][1], single[0][0], hc, rc)


def test_frames_with_tile_blocks_equal_single_frames(vt, vto, ctx):
    """The multi-frame first passes dealt to the XCDs as S x S blocks of 8x8
    tiles (frame regions padded to whole rounds of blocks, image sizes that
    leave partial blocks) give the bytes of single frames with single tiles."""
    from visualtur_amd.prefetch import random_walk

    vol = blobs_volume(64, 14, seed=73)
    sc = Scene(vto, vol, 0.3, 3, 2)
    sc.setup_gpu(vt, ctx, slots=9)
    sizes = [(96, 80), (33, 47), (120, 50), (17, 129), (64, 64)]
    cams = [random_walk(W, H, 64, 1, seed=950 + k, radius=(1.4, 2.2))[0] for k, (W, H) in enumerate(sizes)]
    single = []
    try:
        ctx.set_first_pass_tiles(1)
        for c in cams:
            rg, hg, st = sc.render_gpu(vt, ctx, c)
            single.append((rg, hg))
        for S, max_frames in ((1, 3), (2, 3), (3, 4), (4, 2), (8, 5)):
            print(f"supertile {S} frames in flight {max_frames}", flush=True)
            ctx.set_first_pass_tiles(S)
            ctx.cache_clear()
            out, stats, agg = _frames(vt, ctx, sc, cams, max_frames)
            for k in range(len(cams)):
                assert out[k][1].tobytes() == single[k][1].tobytes(), (S, k)
                assert out[k][0].tobytes() == single[k][0].tobytes(), (S, k)
    finally:
        ctx.set_first_pass_tiles()
        ctx.cache_destroy()


def test_frames_with_mixed_isos_equal_single_frames(vt, vto, ctx):
    """Frames of two isovalues in flight together over one octree: the
    proven sub-block sides are kept for one iso at a time (the newest
    frame's) and rays of the other iso read the sub-block (min, max).  Every
    frame has the bytes of the same frame rendered alone with no sample
    skipped."""
    from visualtur_amd.prefetch import random_walk

    vol = blobs_volume(64, 14, seed=73)
    sc = Scene(vto, vol, 0.3, 3, 2)
    sc.setup_gpu(vt, ctx, slots=9)
    cams = [random_walk(72, 60, 64, 1, seed=950 + k, radius=(1.4, 2.2))[0] for k in range(7)]
    isos = [0.3, 0.45, 0.3, 0.45, 0.45, 0.3, 0.3]
    prms = [vt.render_params(sc.level_o, sc.level_c, iso) for iso in isos]
    ref = []
    ctx.set_march_options(skip_empty=False)
    try:
        for c, p in zip(cams, prms):
            sc.prm_g = p
            rg, hg, _ = sc.render_gpu(vt, ctx, c)
            ref.append((rg, hg))
    finally:
        ctx.set_march_options(skip_empty=True)
    try:
        for max_frames in (1, 3):
            ctx.cache_clear()
            bufs = [(ctx.alloc(c.W * c.H * 16), ctx.alloc(c.W * c.H * 32)) for c in cams]
            with vt.Frames(ctx, max_frames=max_frames, max_rays=72 * 60) as fr:
                tickets = [fr.submit(c, p, *bufs[k]) for k, (c, p) in enumerate(zip(cams, prms))]
                for t in tickets:
                    assert fr.poll(t)[0]
            for k, ((r, h), c) in enumerate(zip(bufs, cams)):
                R = c.W * c.H
                assert h.download(vt.HIT_DTYPE, (R,)).tobytes() == ref[k][1].tobytes(), (max_frames, k)
                assert r.download(np.float32, (R, 4)).tobytes() == ref[k][0].tobytes(), (max_frames, k)
                r.free()
                h.free()
            # and one frame at a time, the sides following each frame's iso
            for k, (c, p) in enumerate(zip(cams, prms)):
                sc.prm_g = p
                rg, hg, _ = sc.render_gpu(vt, ctx, c)
                assert hg.tobytes() == ref[k][1].tobytes() and rg.tobytes() == ref[k][0].tobytes(), k
    finally:
        ctx.cache_destroy()
    assert (ref[1][1]["node"] != 0).mean() > 0.01


def test_frames_refuse_bad_requests(vt, vto, ctx):
    vol = blobs_volume(32, 6, seed=72)
    sc = Scene(vto, vol, 0.3, 3, 2)
    sc.setup_gpu(vt, ctx, slots=8)
    cam = orbit_camera(vt, 32, 32, (16, 16, 16), 60.0, 10.0)
    big = orbit_camera(vt, 64, 64, (16, 16, 16), 60.0, 10.0)
    r, h = ctx.alloc(64 * 64 * 16), ctx.alloc(64 * 64 * 32)
    try:
        for bad in (0, 33):
            with pytest.raises(vt.VTError):
                vt.Frames(ctx, max_frames=bad, max_rays=32 * 32)
        with vt.Frames(ctx, max_frames=2, max_rays=32 * 32) as fr:
            with pytest.raises(vt.VTError):
                fr.submit(big, sc.prm_g, r, h)          # more rays than max_rays
            t = fr.submit(cam, sc.prm_g, r, h)
            assert fr.poll(t)[0]
            with pytest.raises(vt.VTError):
                fr.submit(cam, vt.render_params(2, 2, 0.3), r, h)   # another levelO
            with pytest.raises(vt.VTError):
                fr.poll(t)                                           # a ticket is reported once
    finally:
        r.free()
        h.free()
        ctx.cache_destroy()
