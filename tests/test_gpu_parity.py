"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs."""
import math

import numpy as np
import pytest

from gpu_scenes import Scene, assert_same, bitexact_fraction, blobs_volume, flat_runs, orbit_camera

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def test_camera_rays_bit_exact(vt, vto, ctx):
    for k, (W, H) in enumerate([(1, 1), (7, 5), (64, 48), (257, 131)]):
        cam = orbit_camera(vt, W, H, (32, 32, 32), 90.0, 37.0 * k, 11.0 * k)
        g = ctx.camera_rays(cam)
        c = vto.ray_dirs(cam)
        assert np.array_equal(g.view(np.uint32), c.view(np.uint32)), (W, H)


@pytest.mark.parametrize("n,level_o,seed", [(5, 2, 1), (5, 3, 2), (6, 4, 3), (6, 3, 4)])
def test_trace_nodes_matches_reference_dfs(vt, vto, ctx, n, level_o, seed):
    D = 1 << n
    vol = blobs_volume(D, 8, seed=seed)
    sc = Scene(vto, vol, 0.45, level_o, min(level_o, 2))
    flat, sizes = flat_runs(sc.runs, level_o)
    ctx.octree_upload(flat, sizes, sc.n, level_o)
    K = 24
    for t, theta in enumerate([0.0, 33.0, 141.0, 250.0]):
        cam = orbit_camera(vt, 40, 30, (D / 2, D / 2, D / 2), 2.2 * D, theta, 15.0 + 10 * t)
        log(f"trace n={n} levelO={level_o} theta={theta}")
        nodes, counts = ctx.trace_nodes(cam, level_o, K, 1 << 20)
        dirs = vto.ray_dirs(cam)
        for r in range(cam.W * cam.H):
            exp, cnt = sc.O.traverse(cam.pos, dirs[r], level_o, max_out=K)
            assert counts[r] == cnt, (r, counts[r], cnt)
            assert nodes[r, : min(cnt, K)].tolist() == exp, r


@pytest.mark.parametrize("level_o", [3, 5])
def test_rays_through_node_corners_and_edges(vt, vto, ctx, level_o):
    """Rays whose direction components are equal in pairs or all three
    cross node corners and edges, where children share tnear: the child
    order falls back from the sorting network to the full (tnear, tfar,
    index) order there.  Every emitted node sequence equals the oracle's
    DFS, and frames equal the oracle's."""
    vol = blobs_volume(32, 10, seed=19)
    sc = Scene(vto, vol, 0.3, level_o, 2)
    sc.setup_gpu(vt, ctx, slots=64)
    K = 48
    lines = [((-5.0, -5.0, -5.0), (1.0, 1.0, 1.0)), ((40.0, 40.0, 40.0), (-1.0, -1.0, -1.0)),
             ((-3.0, -3.0, 7.25), (1.0, 1.0, 0.0)), ((16.0, -6.0, -6.0), (0.0, 1.0, 1.0)),
             ((-8.0, 8.0, -8.0), (1.0, 0.0, 1.0)), ((36.0, -4.0, 36.0), (-1.0, 1.0, -1.0)),
             ((8.0, 8.0, 8.0), (1.0, 1.0, 1.0)), ((-2.0, 30.0, -2.0), (1.0, -1.0, 1.0))]
    for pos, look in lines:
        look = np.asarray(look) / np.linalg.norm(look)
        up = np.cross(look, [0.3, 0.5, 0.8])
        up /= np.linalg.norm(up)
        right = np.cross(look, up)
        for W, H in [(1, 1), (5, 5)]:
            cam = vt.make_camera(W, H, pos, look, up, right, fov_v_deg=10.0)
            nodes, counts = ctx.trace_nodes(cam, level_o, K, 1 << 20)
            dirs = vto.ray_dirs(cam)
            for r in range(cam.W * cam.H):
                exp, cnt = sc.O.traverse(cam.pos, dirs[r], level_o, max_out=K)
                assert counts[r] == cnt, (pos, r)
                assert nodes[r, : min(cnt, K)].tolist() == exp, (pos, r)
            rg, hg, _ = sc.render_gpu(vt, ctx, cam)
            rc, hc = sc.render_cpu(vto, cam)
            assert_same(hg, rg, hc, rc)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_tie_prone_rays_walk_like_the_reference(vt, vto, ctx, seed):
    """Seeded cameras at lattice points (origins on node planes, inside and
    outside the cube) looking along directions with equal components: the
    walk meets equal plane times at many nodes, and every node sequence must
    be the oracle's DFS."""
    rng = np.random.default_rng(seed)
    vol = blobs_volume(32, 12, seed=20 + seed)
    sc = Scene(vto, vol, 0.25, 4, 2)
    flat, sizes = flat_runs(sc.runs, 4)
    ctx.octree_upload(flat, sizes, sc.n, 4)
    K = 64
    dirs_set = [(1, 1, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1), (1, 2, 2), (2, 1, 0), (1, 1, 2)]
    for t in range(12):
        pos = rng.integers(-8, 41, 3).astype(np.float64)
        look = np.asarray(dirs_set[int(rng.integers(len(dirs_set)))], np.float64) * rng.choice([-1.0, 1.0], 3)
        look /= np.linalg.norm(look)
        up = np.cross(look, [0.31, 0.57, 0.76])
        up /= np.linalg.norm(up)
        cam = vt.make_camera(9, 9, pos, look, up, np.cross(look, up), fov_v_deg=float(rng.uniform(5, 60)))
        nodes, counts = ctx.trace_nodes(cam, 4, K, 1 << 20)
        dirs = vto.ray_dirs(cam)
        for r in range(cam.W * cam.H):
            exp, cnt = sc.O.traverse(cam.pos, dirs[r], 4, max_out=K)
            assert counts[r] == cnt, (t, r)
            assert nodes[r, : min(cnt, K)].tolist() == exp, (t, r)


def test_degenerate_axis_aligned_rays(vt, vto, ctx):
    """Rays with zero direction components from grid-aligned origins give
    NaN slab distances; the kernel must follow the reference's literal
    bubble sort there."""
    vol = blobs_volume(32, 10, seed=9)
    sc = Scene(vto, vol, 0.3, 3, 2)
    sc.setup_gpu(vt, ctx, slots=64)
    for pos, look, up in [((16, 16, 60), (0, 0, -1), (0, 1, 0)), ((8, 8, 8), (1, 0, 0), (0, 1, 0)),
                          ((16, 40, 16), (0, -1, 0), (0, 0, 1)), ((4, 12, -10), (0, 0, 1), (0, 1, 0))]:
        right = np.cross(np.asarray(look, np.float64), np.asarray(up, np.float64))
        for W, H in [(1, 1), (3, 3), (5, 5)]:
            cam = vt.make_camera(W, H, pos, look, up, right, fov_v_deg=20.0)
            rg, hg, _ = sc.render_gpu(vt, ctx, cam)
            rc, hc = sc.render_cpu(vto, cam)
            assert_same(hg, rg, hc, rc)


@pytest.mark.parametrize("level_o,level_c,inc", [(3, 3, 3), (3, 2, 3), (4, 2, 3), (2, 1, 3)])
def test_render_all_resident(vt, vto, ctx, level_o, level_c, inc):
    vol = blobs_volume(64, 12, seed=level_o * 7 + level_c)
    sc = Scene(vto, vol, 0.25, level_o, level_c)
    sc.setup_gpu(vt, ctx, slots=1 << (3 * level_c), inc=inc)
    for theta in (0.0, 75.0, 200.0):
        cam = orbit_camera(vt, 96, 80, (32, 32, 32), 100.0, theta)
        log(f"render levelO={level_o} levelC={level_c} theta={theta}")
        rg, hg, st = sc.render_gpu(vt, ctx, cam)
        rc, hc = sc.render_cpu(vto, cam)
        assert st.err == 0
        assert_same(hg, rg, hc, rc)
        assert bitexact_fraction(hg, rg, hc, rc) == 1.0
        assert (hc["node"] != 0).mean() > 0.1   # the view actually hits the surface


@pytest.mark.parametrize("slots,defer,late", [(1, 0, 0), (2, 0, 0), (5, 0, 0), (17, 0, 0), (1, 4, 4), (2, 16, 16),
                                              (5, 1, 8), (17, 4, 16), (17, 2, 8), (5, 0, 3)])
def test_capacity_invariance(vt, vto, ctx, slots, defer, late):
    """Hits do not depend on the device cache capacity (the reference's GPU
    tier could overwrite a brick another ray still used, lruCache.cu:254-266)
    nor on how many absent-brick nodes a ray walks past per pass."""
    vol = blobs_volume(64, 12, seed=11)
    sc = Scene(vto, vol, 0.25, 3, 2)
    sc.setup_gpu(vt, ctx, slots=slots)
    cam = orbit_camera(vt, 64, 64, (32, 32, 32), 100.0, 30.0)
    ctx.set_march_options(skip_empty=True, max_defer=defer, max_defer_late=late)
    try:
        rg, hg, st = sc.render_gpu(vt, ctx, cam)
    finally:
        ctx.set_march_options()
    rc, hc = sc.render_cpu(vto, cam)
    assert_same(hg, rg, hc, rc)
    assert st.passes >= 2
    cs = ctx.cache_stats()
    assert cs["gpu_resident"] <= slots


def test_literal_octree_skips_last_octant(vt, vto, ctx):
    """Q1 (CreateOctree.cu:550): with the reference's builder the (+,+,+)
    octant is absent and the GPU follows the file it is given."""
    vol = blobs_volume(32, 12, seed=21)
    sc = Scene(vto, vol, 0.4, 3, 2, literal=True)
    sc.setup_gpu(vt, ctx, slots=64)
    cam = orbit_camera(vt, 64, 64, (16, 16, 16), 70.0, 45.0, 35.0)
    rg, hg, _ = sc.render_gpu(vt, ctx, cam)
    rc, hc = sc.render_cpu(vto, cam)
    assert_same(hg, rg, hc, rc)
    # no hit inside octant 7 = [16,32)^3
    hit = hc["node"] != 0
    v = hc["voxel"][hit]
    assert not np.any((v[:, 0] >= 16) & (v[:, 1] >= 16) & (v[:, 2] >= 16))


def test_hdf5_host_tier_paths(vt, vto, ctx, tmp_path):
    vol = blobs_volume(64, 10, seed=31, dims=(64, 60, 50))
    sc = Scene(vto, vol, 0.5, 3, 2)
    h5 = tmp_path / "vol.h5"
    vt.file_create(h5, "value", vol, chunk=16)
    octf = tmp_path / "vol.oct"
    vto.octree_write(octf, sc.iso, 64, vol.shape, sc.n, sc.runs)
    cam = orbit_camera(vt, 80, 64, (32, 30, 25), 140.0, 120.0)
    rc, hc = sc.render_cpu(vto, cam)
    for cpu_slots, gpu_slots in [(64, 64), (64, 6), (9, 5)]:
        log(f"hdf5 cpu_slots={cpu_slots} gpu_slots={gpu_slots}")
        iso, n, rd = ctx.octree_load_file(octf, 3)
        assert (n, rd) == (6, (64, 60, 50)) and iso == np.float32(0.5)
        ctx.cache_create(2, 16, 3, gpu_slots, cpu_slots)
        ctx.cache_attach_hdf5(h5, "value")
        if cpu_slots >= 64:
            ctx.cache_prefetch_host(wait=True)
        prm = vt.render_params(3, 2, iso)
        R = cam.W * cam.H
        d_rgba, d_hits = ctx.alloc(R * 16), ctx.alloc(R * 32)
        ctx.render_frame(cam, prm, d_rgba, d_hits)
        rg, hg = d_rgba.download(np.float32, (R, 4)), d_hits.download(vt.HIT_DTYPE, (R,))
        d_rgba.free(); d_hits.free()
        assert_same(hg, rg, hc, rc)


def test_file_read_box_matches_oracle(vt, vto, tmp_path):
    rng = np.random.default_rng(4)
    vol = rng.standard_normal((20, 18, 16)).astype(np.float32)
    p = tmp_path / "r.h5"
    vt.file_create(p, "value", vol, chunk=8)
    V = vto.Volume(vol)
    with vt.File(p, "value") as f:
        assert f.dims == (20, 18, 16)
        for s, e in [((0, 0, 0), (10, 10, 10)), ((-3, -2, -1), (7, 8, 9)), ((15, 15, 10), (25, 25, 20)),
                     ((30, 0, 0), (34, 4, 4)), ((-8, -8, -8), (-1, 3, 3))]:
            assert np.array_equal(f.read_box(s, e), V.read_box(s, e))


def test_error_paths(vt, vto, ctx, tmp_path):
    c = vt.Context(0)
    try:
        cam = vt.make_camera(8, 8, (0, 0, 50))
        prm = vt.render_params(2, 1, 0.5)
        buf = c.alloc(64 * 16)
        with pytest.raises(vt.VTError) as e:
            c.render_frame(cam, prm, buf, None)
        assert e.value.code == -5   # VT_E_STATE: no octree
        bad = np.array([9, 8], np.uint64)   # start after end
        with pytest.raises(vt.VTError) as e:
            c.octree_upload(np.concatenate([np.array([1, 1], np.uint64), bad]), [2, 2], 3, 1)
        assert e.value.code == -4   # VT_E_FORMAT
        f = tmp_path / "bad.oct"
        f.write_bytes(b"\x01\x02\x03\x04" + b"\x00" * 40)
        with pytest.raises(vt.VTError) as e:
            c.octree_load_file(f, 2)
        assert e.value.code == -4
        with pytest.raises(vt.VTError) as e:
            vt.File(tmp_path / "missing.h5", "value")
        assert e.value.code == -3   # VT_E_IO
        buf.free()
    finally:
        c.close()


def test_halo_too_small_is_reported(vt, vto, ctx):
    """With cubeInc=0 the march reads past the brick's upper faces (the
    reference reads out of its slot, Q4).  Every read is bounds-checked:
    the frame returns VT_E_RANGE instead of reading outside the slot."""
    vol = blobs_volume(64, 12, seed=11)
    sc = Scene(vto, vol, 0.25, 3, 2)
    sc.setup_gpu(vt, ctx, slots=64, inc=0)
    cam = orbit_camera(vt, 64, 64, (32, 32, 32), 100.0, 30.0)
    with pytest.raises(vt.VTError) as e:
        sc.render_gpu(vt, ctx, cam)
    assert e.value.code == -7


def test_c1_full_frame(vt, vto, ctx):
    """Config 1: 256^3, 16 blobs, iso 0.5, 32^3 bricks (levels 0..3, all 512
    resident), one 512x512 frame, 45 deg fov, camera 2.5 extents from the
    centre looking at it."""
    vol = blobs_volume(256, 16, seed=0xC0FFEE)
    sc = Scene(vto, vol, 0.5, 3, 3)
    sc.setup_gpu(vt, ctx, slots=512)
    cam = vt.make_camera(512, 512, (128.0, 128.0, 128.0 + 2.5 * 256.0), fov_v_deg=45.0)
    log("c1 gpu")
    rg, hg, st = sc.render_gpu(vt, ctx, cam)
    log(f"c1 gpu done passes={st.passes}; oracle")
    rc, hc = sc.render_cpu(vto, cam)
    assert_same(hg, rg, hc, rc)
    assert bitexact_fraction(hg, rg, hc, rc) == 1.0
    assert (hc["node"] != 0).mean() > 0.02   # the 16 blobs at iso 0.5 cover ~5% of the frame


@pytest.mark.parametrize("iso,seed", [(0.25, 41), (0.6, 42), (0.05, 43)])
def test_sample_skipping_is_exact(vt, vto, ctx, iso, seed):
    """The sub-block skip evaluates fewer samples but gives bit-identical
    frames (and equals the oracle)."""
    vol = blobs_volume(64, 14, seed=seed)
    sc = Scene(vto, vol, iso, 3, 2)
    sc.setup_gpu(vt, ctx, slots=64)
    cam = orbit_camera(vt, 96, 72, (32, 32, 32), 100.0, 20.0 * seed)
    ctx.set_frame_options(collect_stats=True)
    try:
        ctx.set_march_options(skip_empty=False)
        r0, h0, s0 = sc.render_gpu(vt, ctx, cam)
        ctx.set_march_options(skip_empty=True)
        r1, h1, s1 = sc.render_gpu(vt, ctx, cam)
    finally:
        ctx.set_march_options(skip_empty=True)
        ctx.set_frame_options(False, False)
    assert np.array_equal(h0, h1)
    assert np.array_equal(r0.view(np.uint32), r1.view(np.uint32))
    rc, hc = sc.render_cpu(vto, cam)
    assert_same(h1, r1, hc, rc)
    assert s1.samples < s0.samples
    print(f"samples without skip {s0.samples}, with skip {s1.samples}", flush=True)


@pytest.mark.parametrize("level_o,slots", [(3, 40), (4, 64), (4, 512)])
def test_deferred_nodes_are_exact(vt, vto, ctx, level_o, slots):
    """Walking past absent bricks (marching them in a later pass, before the
    tentative result is accepted) changes the pass count, never a byte of the
    frame.  512 slots hold every brick: then only the cold start defers."""
    vol = blobs_volume(64, 16, seed=51)
    sc = Scene(vto, vol, 0.3, level_o, 3)
    sc.setup_gpu(vt, ctx, slots=slots)
    ctx.set_frame_options(collect_stats=True)
    ctx.set_early_loads(0)   # pass counts are compared: no timing-dependent loads
    try:
        for theta in (10.0, 140.0):
            cam = orbit_camera(vt, 96, 80, (32, 32, 32), 110.0, theta, 25.0)
            rc, hc = sc.render_cpu(vto, cam)
            ref, passes = None, {}
            for defer in (0, 1, 4, 16):
                log(f"defer={defer} levelO={level_o} slots={slots} theta={theta}")
                ctx.set_march_options(skip_empty=True, max_defer=defer)
                ctx.cache_clear()
                rg, hg, st = sc.render_gpu(vt, ctx, cam)
                assert st.err == 0
                passes[defer] = st.passes
                if ref is None:
                    assert_same(hg, rg, hc, rc)
                    ref = (hg.tobytes(), rg.tobytes())
                else:
                    assert (hg.tobytes(), rg.tobytes()) == ref, defer
            print(f"passes by max_defer: {passes}", flush=True)
            if slots == 512:   # room for every request: deferral saves passes
                assert passes[16] < passes[0]
    finally:
        ctx.set_march_options()
        ctx.set_frame_options(False, False)
        ctx.set_early_loads()


@pytest.mark.parametrize("slots,defer", [(512, 2), (5, 0), (17, 2), (40, 4)])
def test_persistent_pass_matches_tiled_pass(vt, vto, ctx, slots, defer):
    """The persistent pass (lanes refilled with new rays) and the one-lane-
    per-ray tiled pass give the same bytes, pass count and requests."""
    vol = blobs_volume(64, 16, seed=71)
    sc = Scene(vto, vol, 0.3, 4, 3)
    sc.setup_gpu(vt, ctx, slots=slots)
    cams = [orbit_camera(vt, 131, 77, (32, 32, 32), 105.0, 30.0 + 50.0 * k, 15.0) for k in range(3)]
    outs = {}
    ctx.set_early_loads(0)   # the persistent pass loads between passes only
    try:
        for persistent in (False, True):
            log(f"persistent={persistent} slots={slots} defer={defer}")
            ctx.set_march_options(skip_empty=True, max_defer=defer, persistent=persistent)
            ctx.cache_clear()
            rows = []
            for cam in cams:
                rg, hg, st = sc.render_gpu(vt, ctx, cam)
                assert st.err == 0
                rows.append((hg.tobytes(), rg.tobytes(), st.passes, st.bricks_requested))
            outs[persistent] = rows
    finally:
        ctx.set_march_options()
        ctx.set_early_loads()
    assert outs[False] == outs[True]
    rc, hc = sc.render_cpu(vto, cams[-1])
    assert_same(np.frombuffer(outs[True][-1][0], vt.HIT_DTYPE), np.frombuffer(outs[True][-1][1], np.float32).reshape(-1, 4),
                hc, rc)


@pytest.mark.parametrize("slots,defer,skip", [(5, 0, True), (17, 2, True), (40, 4, True), (17, 2, False),
                                              (512, 2, True)])
def test_wide_march_matches_one_lane(vt, vto, ctx, slots, defer, skip):
    """Passes with 2..64 lanes per ray (the later passes at every fixed
    width and by the lane budget, and the first pass too) give the bytes,
    pass count, requests, sample and node counts of the one-lane passes."""
    vol = blobs_volume(64, 16, seed=91)
    sc = Scene(vto, vol, 0.3, 4, 3)
    sc.setup_gpu(vt, ctx, slots=slots)
    cams = [orbit_camera(vt, 131, 77, (32, 32, 32), 105.0, 20.0 + 70.0 * k, 15.0) for k in range(3)]
    outs = {}
    ctx.set_frame_options(collect_stats=True)
    ctx.set_early_loads(0)   # loads beside the first pass depend on timing, and so would the counts
    try:
        for wide in (0, -1, -2, -3, -4, -5, -6, 4096, "first-2", "first-6", "first"):
            log(f"wide={wide} slots={slots} defer={defer} skip={skip}")
            ctx.set_march_options(skip_empty=skip, max_defer=defer)
            if wide == "first-2":     # 4 lanes per ray in every pass
                ctx.set_wide_march(-2, -2)
            elif wide == "first-6":   # one wave per ray in every pass
                ctx.set_wide_march(-6, -6)
            elif wide == "first":     # budgets: 2 lanes per ray in the first pass of 131x77
                ctx.set_wide_march(1 << 18, 2 * 131 * 77)
            else:
                ctx.set_wide_march(wide, 0)
            ctx.cache_clear()
            rows = []
            for cam in cams:
                rg, hg, st = sc.render_gpu(vt, ctx, cam)
                assert st.err == 0
                rows.append((hg.tobytes(), rg.tobytes(), st.passes, st.bricks_requested, st.samples, st.nodes))
            outs[wide] = rows
    finally:
        ctx.set_wide_march()
        ctx.set_march_options()
        ctx.set_frame_options(False, False)
        ctx.set_early_loads()
    for wide, rows in outs.items():
        for k, (row, ref) in enumerate(zip(rows, outs[0])):
            assert row[:2] == ref[:2], (wide, k, "bytes")
            assert row[2:] == ref[2:], (wide, k, row[2:], ref[2:])
    if slots < 512:
        assert max(r[2] for r in outs[0]) >= 2   # later passes ran
    rc, hc = sc.render_cpu(vto, cams[-1])
    assert_same(np.frombuffer(outs[-6][-1][0], vt.HIT_DTYPE), np.frombuffer(outs[-6][-1][1], np.float32).reshape(-1, 4),
                hc, rc)


@pytest.mark.parametrize("slots,reserve", [(17, 4), (40, 16), (200, 64), (200, 512)])
def test_early_loads_are_exact(vt, vto, ctx, slots, reserve):
    """Copying requested bricks beside the first pass (into a reserve of
    free slots, published only between passes) never changes a byte of the
    frame."""
    vol = blobs_volume(64, 16, seed=101)
    sc = Scene(vto, vol, 0.3, 4, 3)
    sc.setup_gpu(vt, ctx, slots=slots)
    cams = [orbit_camera(vt, 320, 240, (32, 32, 32), 95.0, 15.0 + 40.0 * k, 20.0) for k in range(4)]
    outs, early = {}, 0
    try:
        for res in (0, reserve):
            log(f"early reserve={res} slots={slots}")
            ctx.set_early_loads(res)
            ctx.cache_clear()
            rows = []
            for cam in cams:
                rg, hg, st = sc.render_gpu(vt, ctx, cam)
                assert st.err == 0
                rows.append((hg.tobytes(), rg.tobytes()))
                if res:
                    early += st.bricks_early
            outs[res] = rows
    finally:
        ctx.set_early_loads()
    assert outs[reserve] == outs[0]
    print(f"bricks loaded beside the first pass: {early}", flush=True)
    rc, hc = sc.render_cpu(vto, cams[-1])
    assert_same(np.frombuffer(outs[reserve][-1][0], vt.HIT_DTYPE),
                np.frombuffer(outs[reserve][-1][1], np.float32).reshape(-1, 4), hc, rc)


def test_early_loads_from_the_host_tier(vt, vto, ctx, tmp_path):
    """Early loads from the pinned host tier, filled ahead and on demand."""
    vol = blobs_volume(64, 10, seed=33, dims=(64, 60, 50))
    sc = Scene(vto, vol, 0.5, 3, 2)
    h5 = tmp_path / "vol.h5"
    vt.file_create(h5, "value", vol, chunk=16)
    cams = [orbit_camera(vt, 256, 200, (32, 30, 25), 120.0, 60.0 * k) for k in range(3)]
    prm = vt.render_params(3, 2, 0.5)
    flat, sizes = flat_runs(sc.runs, 3)
    for cpu_slots, gpu_slots, reserve in [(64, 12, 6), (9, 6, 3)]:
        outs = {}
        for res in (0, reserve):
            log(f"host tier cpu={cpu_slots} gpu={gpu_slots} reserve={res}")
            ctx.octree_upload(flat, sizes, sc.n, 3)
            ctx.cache_create(2, 16, 3, gpu_slots, cpu_slots)
            ctx.cache_attach_hdf5(h5, "value")
            if cpu_slots >= 64:
                ctx.cache_prefetch_host(wait=True)
            ctx.set_early_loads(res)
            try:
                rows = []
                for cam in cams:
                    R = cam.W * cam.H
                    d_rgba, d_hits = ctx.alloc(R * 16), ctx.alloc(R * 32)
                    st = ctx.render_frame(cam, prm, d_rgba, d_hits)
                    assert st.err == 0
                    rows.append((d_hits.download(vt.HIT_DTYPE, (R,)).tobytes(),
                                 d_rgba.download(np.float32, (R, 4)).tobytes()))
                    d_rgba.free(); d_hits.free()
                outs[res] = rows
            finally:
                ctx.set_early_loads()
        assert outs[reserve] == outs[0]
    rc, hc = sc.render_cpu(vto, cams[-1])
    R = cams[-1].W * cams[-1].H
    assert_same(np.frombuffer(outs[reserve][-1][0], vt.HIT_DTYPE),
                np.frombuffer(outs[reserve][-1][1], np.float32).reshape(R, 4), hc, rc)


@pytest.mark.parametrize("slots", [512, 17])
def test_first_pass_budget_is_exact(vt, vto, ctx, slots):
    """Rays that stop after a few walk nodes in the first pass and finish in
    later passes (wide or not) give the same bytes; with every brick
    resident the budget alone adds the passes."""
    vol = blobs_volume(64, 16, seed=111)
    sc = Scene(vto, vol, 0.3, 4, 3)
    sc.setup_gpu(vt, ctx, slots=slots)
    cams = [orbit_camera(vt, 160, 120, (32, 32, 32), 100.0, 25.0 + 60.0 * k, 15.0) for k in range(3)]
    outs, passes = {}, {}
    try:
        for budget, wide in ((0, 1 << 18), (1, 1 << 18), (2, 0), (4, 1 << 18), (16, 1 << 18)):
            log(f"budget={budget} wide={wide} slots={slots}")
            ctx.set_first_pass_budget(budget)
            ctx.set_wide_march(wide)
            ctx.cache_clear()
            rows, ps = [], []
            for cam in cams + cams[-1:]:   # the last view twice: the second time every brick it needs is resident
                rg, hg, st = sc.render_gpu(vt, ctx, cam)
                assert st.err == 0
                rows.append((hg.tobytes(), rg.tobytes()))
                ps.append(st.passes)
            outs[budget], passes[budget] = rows, ps
    finally:
        ctx.set_first_pass_budget()
        ctx.set_wide_march()
    for budget, rows in outs.items():
        assert rows == outs[0], budget
    print(f"passes by budget: {passes}", flush=True)
    if slots == 512:   # the repeated view is resident: only the budget splits the walk
        assert passes[0][-1] == 1 and passes[1][-1] >= 2
    rc, hc = sc.render_cpu(vto, cams[-1])
    assert_same(np.frombuffer(outs[1][-1][0], vt.HIT_DTYPE), np.frombuffer(outs[1][-1][1], np.float32).reshape(-1, 4),
                hc, rc)


def test_iso_change_keeps_the_cache_exact(vt, vto, ctx):
    """The sub-block side maps follow the frame's isovalue: a cache filled
    at one iso renders a second iso (with its own octree) exactly."""
    vol = blobs_volume(64, 14, seed=81)
    a, b = Scene(vto, vol, 0.25, 4, 2), Scene(vto, vol, 0.45, 4, 2)
    a.setup_gpu(vt, ctx, slots=40)
    b.prm_g = vt.render_params(b.level_o, b.level_c, b.iso)
    cam = orbit_camera(vt, 96, 72, (32, 32, 32), 100.0, 65.0)
    for sc in (a, b, a):
        flat, sizes = flat_runs(sc.runs, sc.level_o)
        ctx.octree_upload(flat, sizes, sc.n, sc.level_o)
        log(f"iso {sc.iso}")
        rg, hg, st = sc.render_gpu(vt, ctx, cam)   # the cache keeps the bricks of the previous iso
        rc, hc = sc.render_cpu(vto, cam)
        assert st.err == 0
        assert_same(hg, rg, hc, rc)


@pytest.mark.parametrize("slots", [512, 17])
def test_first_pass_tile_blocks_are_exact(vt, vto, ctx, slots):
    """The one-lane first pass gives the same bytes whether the XCDs are
    dealt single 8x8 tiles or S x S blocks of them (image sizes that leave
    partial blocks at the right and bottom edges)."""
    vol = blobs_volume(64, 16, seed=111)
    sc = Scene(vto, vol, 0.3, 4, 3)
    sc.setup_gpu(vt, ctx, slots=slots)
    cams = [orbit_camera(vt, 160, 120, (32, 32, 32), 100.0, 25.0 + 60.0 * k, 15.0) for k in range(2)]
    cams.append(orbit_camera(vt, 93, 71, (32, 32, 32), 90.0, 200.0, -10.0))
    outs = {}
    try:
        ctx.set_wide_march(1 << 18, 0)   # the first pass runs one lane per ray
        for S in (1, 2, 3, 4, 8):
            log(f"supertile={S} slots={slots}")
            ctx.set_first_pass_tiles(S)
            ctx.cache_clear()
            rows = []
            for cam in cams:
                rg, hg, st = sc.render_gpu(vt, ctx, cam)
                assert st.err == 0
                rows.append((hg.tobytes(), rg.tobytes()))
            outs[S] = rows
    finally:
        ctx.set_first_pass_tiles()
        ctx.set_wide_march()
    for S, rows in outs.items():
        assert rows == outs[1], S
    rc, hc = sc.render_cpu(vto, cams[-1])
    assert_same(np.frombuffer(outs[4][-1][0], vt.HIT_DTYPE), np.frombuffer(outs[4][-1][1], np.float32).reshape(-1, 4),
                hc, rc)
