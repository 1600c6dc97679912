#!/usr/bin/env python3
"""Benchmark of the isosurface ray-casting hot path on one MI355X.

One step = one rendered frame (every primary ray of the frame traversed,
marched and shaded, all passes and brick loads included).  Prints one JSON
line: rays/sec over the timed frames, p50/p99 frame ms, the roofline of the
render-pass kernel, a bounded CPU baseline of the reference algorithm, and a
parity check of sampled rays against the CPU oracle.

Workloads
  c3  (default) 2048^3 signed-distance field of 200 seeded spheres and tori
      plus 0.02-amplitude fBm, iso 0, written to HDF5 in 32^3 chunks; 32^3
      bricks with a 3-voxel halo (levels 0..6, 262144 bricks); device brick
      cache 8 GB, pinned host tier 32 GB filled from the file before timing;
      1920x1080 frames along a seeded random-walk camera path.  Legs: plain,
      and with the learned prefetcher; each leg's sampled frames must equal
      the plain leg's bytes.
  c2  1024^3 value-noise fBm (6 octaves, lacunarity 2, gain 0.5), iso 0,
      HDF5 in 32^3 chunks; device cache capped at 25% (8192 of 32768
      bricks), pinned host tier; 1920x1080, orbit at 2.2 extents from the
      centre, 5.6 deg per frame.
  c1  256^3 Gaussian blobs (16, seeded), iso 0.5, 32^3 bricks (levels 0..3,
      all 512 resident), 512x512, 45 deg fov, camera 2.5 extents from the
      centre looking at it.
  c4  serving: the c1 and c2 volumes resident, Poisson request streams.
  c5  4096x4096x2048 turbulence, 3840x2160 fly-through.
"""
from __future__ import annotations

import argparse
import json
import shutil
import tempfile
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import visualtur_amd as vt  # noqa: E402
from visualtur_amd.volumes import C1_SEED, gaussian_blobs  # noqa: E402

C2_SEED = 0x5EED0002
METRIC = "rays/sec + p50/p99 frame ms on 1 MI355X, prefetch vs plain, identical hits"
HBM_PEAK_GBS = 8000.0


# rocprofv3 summaries of the render passes (k_render_pass, one lane per ray, and
# k_render_wide, 2^k lanes per ray) for the workloads: FETCH_SIZE + WRITE_SIZE per
# dispatch, each scaled by what the same counter reads for a 256 MiB copy
PROFILES = {"c1": ROOT / "profiles" / "render_passes_256cubed_blobs.md",
            "c2": ROOT / "profiles" / "render_passes_1024cubed_fbm_orbit.md",
            "c3": ROOT / "profiles" / "render_passes_2048cubed_sdf_walk.md",
            "c5": ROOT / "profiles" / "render_passes_4096x4096x2048_turbulence_flythrough.md"}
RENDER_KERNELS = "k_render_pass+k_render_wide"


def measured_traffic(name: str):
    """HBM bytes per render-pass launch from the committed rocprof summary, or None."""
    import re
    path = PROFILES.get(name)
    if path is None or not path.exists():
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", path.read_text(), re.M)
    return int(m.group(1)) if m else None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu() -> dict:
    """CPU model, logical CPUs of the machine and CPUs this process may use."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "usable": len(os.sched_getaffinity(0))}


def march_options(ctx, args) -> None:
    ctx.set_march_options(skip_empty=True, max_defer=args.max_defer, persistent=bool(args.persistent),
                          max_defer_late=None if args.max_defer_late < 0 else args.max_defer_late)


def trace_bricks(ctx, args, cams, prm, level_c, d_rgba, d_hits, collect_stats=False):
    """Distinct bricks marched per frame under the reference's pass protocol
    (and, with collect_stats, the trilinear samples per frame): a replay with
    no deferral (a ray parks at its first absent brick), whose visited-brick
    trace equals the set the reference algorithm marches.  The timed runs'
    march options are restored afterwards."""
    ctx.set_frame_options(collect_stats=collect_stats, record_trace=True)
    ctx.set_march_options(skip_empty=True, max_defer=0, max_defer_late=0)
    U, samples = [], []
    try:
        for cam in cams:
            st = ctx.render_frame(cam, prm, d_rgba, d_hits)
            U.append(int(ctx.brick_trace(level_c).sum()))
            samples.append(int(st.samples))
    finally:
        ctx.set_frame_options(False, False)
        march_options(ctx, args)
    return (U, samples) if collect_stats else U


def brick_bytes(cube: int, inc: int) -> int:
    return (cube + 2 * inc) ** 3 * 4


class Workload:
    pass


def setup_c1(ctx, args) -> Workload:
    w = Workload()
    w.name = "c1"
    w.desc = "256^3 gaussian blobs x16, iso 0.5, 32^3 bricks all resident, 512x512, 45deg, d=2.5 extents"
    w.vol = gaussian_blobs(256, 16, C1_SEED)
    w.iso, w.level_o, w.level_c, w.inc = 0.5, 3, 3, 3
    w.d_vol = ctx.to_device(w.vol)
    t0 = time.perf_counter()
    w.n, sizes = ctx.octree_build(w.d_vol, w.vol.shape, w.iso, False, w.level_o)
    w.runs = [ctx.octree_runs(l, sizes[l]) for l in range(w.level_o + 1)]
    log(f"octree built on the GPU in {time.perf_counter() - t0:.2f}s (n={w.n})")
    w.cube = 1 << (w.n - w.level_c)
    w.slots = 1 << (3 * w.level_c)
    ctx.cache_create(w.level_c, w.cube, w.inc, w.slots, 0)
    ctx.cache_attach_device_volume(w.d_vol, w.vol.shape)
    w.W, w.H = 512, 512
    cam = vt.make_camera(w.W, w.H, (128.0, 128.0, 128.0 + 2.5 * 256.0), fov_v_deg=45.0)
    w.cameras = lambda k: cam
    w.prm = vt.render_params(w.level_o, w.level_c, w.iso)
    w.oracle_volume = lambda: w.vol
    w.oracle_runs = lambda: (w.n, w.runs)
    w.extra = {"volume": "256^3", "bricks": 512, "cache_slots": 512, "frame": "512x512"}
    return w


def orbit(vt_mod, W, H, center, radius, deg, elev_deg=0.0, fov=45.0):
    t, e = math.radians(deg), math.radians(elev_deg)
    eye = np.array(center, np.float64) + radius * np.array([math.sin(t) * math.cos(e), math.sin(e),
                                                            math.cos(t) * math.cos(e)])
    look, up, right = vt_mod.look_at(eye, center)
    return vt_mod.make_camera(W, H, eye, look, up, right, fov_v_deg=fov)


def setup_c2(ctx, args) -> Workload:
    w = Workload()
    w.name = "c2"
    w.desc = ("1024^3 value-noise fBm (6 oct, lac 2, gain 0.5), iso 0, HDF5 of 32^3 bricks with a 3-voxel halo "
              "(and the 32^3-chunked volume for the CPU baseline), device cache 25% (8192 of 32768 bricks), pinned "
              "host tier, 1920x1080, 45deg, orbit r=2.2 extents, 5.6 deg/frame")
    dims = (1024, 1024, 1024)
    w.iso, w.level_o, w.level_c, w.inc = 0.0, 5, 5, 3
    t0 = time.perf_counter()
    d_vol = ctx.alloc(4 * 1024 ** 3)
    ctx.volume_generate(vt.GEN_FBM, dims, C2_SEED, d_vol, octaves=6, base_freq=1.0 / 64.0, gain=0.5,
                        lacunarity=2.0, scale=1.0)
    log(f"volume generated in {time.perf_counter() - t0:.2f}s")
    t0 = time.perf_counter()
    w.n, sizes = ctx.octree_build(d_vol, dims, w.iso, False, w.level_o)
    w.runs = [ctx.octree_runs(l, sizes[l]) for l in range(w.level_o + 1)]
    log(f"octree built on the GPU in {time.perf_counter() - t0:.2f}s (n={w.n}, leaf runs={sizes[-1]})")
    w.vol = d_vol.download(np.float32, dims) if not args.no_cpu else None
    w.tmpdir = tempfile.mkdtemp(prefix="vt_c2_")
    h5 = os.path.join(w.tmpdir, "c2.h5")
    t0 = time.perf_counter()
    ctx.volume_write_hdf5(d_vol, dims, h5, "value", 32)
    log(f"hdf5 written in {time.perf_counter() - t0:.2f}s")
    # the same bricks as a brick-major file (32^3 bricks with their halo, one
    # record each): the host tier's source unless --c2-source volume
    from visualtur_amd import prefetch as pf

    codes = pf.candidate_codes(w.runs[w.level_c])
    h5b = os.path.join(w.tmpdir, "c2_bricks.h5")
    t0 = time.perf_counter()
    ctx.volume_write_bricks_hdf5(d_vol, dims, h5b, w.level_c, 1 << (w.n - w.level_c), w.inc, codes)
    log(f"brick-major hdf5 written in {time.perf_counter() - t0:.2f}s ({codes.size} bricks)")
    d_vol.free()
    w.cube = 1 << (w.n - w.level_c)
    nbricks = 1 << (3 * w.level_c)
    w.slots = nbricks // 4
    ctx.cache_create(w.level_c, w.cube, w.inc, w.slots, nbricks)
    if args.c2_source == "bricks":
        ctx.cache_attach_hdf5_bricks(h5b)
    else:
        ctx.cache_attach_hdf5(h5, "value")
    t0 = time.perf_counter()
    ctx.cache_prefetch_host(wait=True)
    fill_s = time.perf_counter() - t0
    log(f"host tier filled from the {args.c2_source} file in {fill_s:.2f}s")
    w.W, w.H = 1920, 1080
    w.cameras = lambda k: orbit(vt, w.W, w.H, (512.0, 512.0, 512.0), 2.2 * 1024.0, 5.6 * k)
    w.prm = vt.render_params(w.level_o, w.level_c, w.iso)
    w.oracle_volume = lambda: w.vol
    w.oracle_runs = lambda: (w.n, w.runs)
    w.cpu_cache = (h5, w.slots)   # the CPU baseline reads the same file through an LRU of the same capacity
    w.extra = {"volume": "1024^3", "bricks": nbricks, "cache_slots": w.slots, "host_tier_slots": nbricks,
               "host_tier_source": f"{args.c2_source} file ({'one record per brick with halo' if args.c2_source == 'bricks' else '32^3 chunks'}), filled in {fill_s:.2f}s before timing",
               "frame": "1920x1080", "frames": "orbit"}
    return w


def prefetch_leg(ctx, w, cams, args, d_rgba, d_hits, plain_ms, plain_uploads, U, plain_passes):
    """Learned prefetching against the plain run just timed: train the MLP on
    the visited-brick traces of other camera paths, clear the device cache,
    replay the same frames with a prefetch before each, and compare."""
    from visualtur_amd import prefetch as pf

    t0 = time.perf_counter()
    cand = pf.candidate_codes(w.runs[w.level_c])
    tcams = [c for s in range(args.train_paths)
             for c in pf.random_walk(w.W // 4, w.H // 4, 1 << w.n, args.train_frames, seed=9000 + s,
                                     radius=(1.9, 2.6))]
    traces = pf.collect_traces(ctx, tcams, w.prm, w.level_c)
    net = vt.MLP(ctx, seed=11, max_batch=65536)
    X, Y = pf.build_dataset(ctx, net, tcams, traces, cand, w.n, w.level_c, neg_per_pos=3, seed=2)
    losses, pos_w = pf.train(ctx, net, X, Y, steps=args.train_steps, batch=8192, lr=1e-3)
    train_s = time.perf_counter() - t0
    log(f"prefetcher trained in {train_s:.1f}s on {len(tcams)} frames / {X.shape[0]} samples, "
        f"loss {losses[:10].mean():.3f} -> {losses[-10:].mean():.3f}")
    R = w.W * w.H
    ctx.cache_clear()
    for k in range(args.warmup):
        net.prefetch(cams[k], w.slots, args.prefetch_min_prob)
        ctx.render_frame(cams[k], w.prm, d_rgba, d_hits)
    ctx.synchronize()
    ms, pre_ms, uploads, pre_loaded, passes = [], [], [], [], []
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        f0 = time.perf_counter()
        pre_loaded.append(net.prefetch(cams[k], w.slots, args.prefetch_min_prob))
        f1 = time.perf_counter()
        st = ctx.render_frame(cams[k], w.prm, d_rgba, d_hits)
        f2 = time.perf_counter()
        ms.append((f2 - f0) * 1e3)
        pre_ms.append((f1 - f0) * 1e3)
        uploads.append(st.bricks_uploaded)
        passes.append(st.passes)
    ctx.synchronize()
    total = time.perf_counter() - t0
    # the last frame must be identical to the plain render of the same camera
    hits_pf = d_hits.download(vt.HIT_DTYPE, (R,))
    rgba_pf = d_rgba.download(np.float32, (R, 4))
    ctx.cache_clear()
    ctx.render_frame(cams[args.warmup + args.steps - 1], w.prm, d_rgba, d_hits)
    hits_pl = d_hits.download(vt.HIT_DTYPE, (R,))
    rgba_pl = d_rgba.download(np.float32, (R, 4))
    # bytewise: a 0/0 secant gives NaN depths, which must match as NaN
    differ = int(np.count_nonzero((hits_pf.view(np.uint8).reshape(R, -1) != hits_pl.view(np.uint8).reshape(R, -1)).any(1)
                                  | (rgba_pf.view(np.uint32) != rgba_pl.view(np.uint32)).any(1)))
    same = differ == 0
    net.close()
    m = np.array(ms)
    return {"value": round(R * args.steps / total, 1), "unit": "rays/s",
            "p50_ms": round(float(np.percentile(m, 50)), 4), "p99_ms": round(float(np.percentile(m, 99)), 4),
            "prefetch_ms_per_frame": round(float(np.mean(pre_ms)), 4), "candidates": int(cand.size),
            "demand_misses_per_frame": float(np.mean(uploads)),
            "prefetched_per_frame": float(np.mean(pre_loaded)),
            "device_miss_rate": round(float(sum(uploads)) / max(1, sum(U)), 5),
            "plain_device_miss_rate": round(float(sum(plain_uploads)) / max(1, sum(U)), 5),
            "passes_per_frame": float(np.mean(passes)), "plain_passes_per_frame": float(np.mean(plain_passes)),
            "plain_value": round(R * len(plain_ms) / (sum(plain_ms) * 1e-3), 1),
            "identical_last_frame": same, "rays_differing": differ, "train_seconds": round(train_s, 2),
            "train_frames": len(tcams), "train_samples": int(X.shape[0]),
            "final_loss": round(float(losses[-10:].mean()), 4)}


def serving_sweep(ctx, args):
    """Config 4, the serving form: the volumes of the out-of-core
    configurations resident as two scenes, the 1024^3 value-noise fBm (iso
    uniform over {-0.2, 0, 0.2}) and the 2048^3 SDF (iso 0); 32 viewers on
    seeded random walks issue requests as a Poisson stream; resolution
    uniform over {256^2, 512^2, 1024^2, 1920x1080}.  Batched serving (mode 0)
    against one request per launch (mode 1) and against the plain renderer
    (vt_render_frame, one request at a time in arrival order), per arrival
    rate."""
    from visualtur_amd import prefetch as pf
    from visualtur_amd.prefetch import random_walk
    from visualtur_amd.volumes import C3_SEED, sdf_shapes

    sv = vt.Server(ctx, hbm_budget=4 << 30, max_tiles=65536)
    DA, DB = 1024, 2048
    dA = ctx.alloc(4 * DA ** 3)
    ctx.volume_generate(vt.GEN_FBM, (DA,) * 3, C2_SEED, dA, octaves=6, base_freq=1.0 / 64.0, gain=0.5,
                        lacunarity=2.0, scale=1.0)
    dB = ctx.alloc(4 * DB ** 3)
    ctx.volume_generate(vt.GEN_SDF, (DB,) * 3, C3_SEED, dB, shapes=sdf_shapes(200, C3_SEED, DB), n=DB, octaves=5,
                        base_freq=1.0 / 32.0, gain=0.5, lacunarity=2.0, scale=0.02)
    t0 = time.perf_counter()
    sA, nA = sv.add_scene_range(dA, (DA,) * 3, 5, 5, -0.2, 0.2)
    sB, nB = sv.add_scene_range(dB, (DB,) * 3, 6, 6, 0.0, 0.0)
    log(f"serving scenes resident in {time.perf_counter() - t0:.1f}s: fBm {nA} bricks, SDF {nB} bricks")
    scenes = {sA: (dA, (DA,) * 3, 5, (1.6, 2.6), (-0.2, 0.0, 0.2)), sB: (dB, (DB,) * 3, 6, (1.3, 2.2), (0.0,))}
    res_set = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]
    # the plain renderer per request: one context per (scene, iso), its octree
    # built for that iso and every brick holding nodes loaded up front (the
    # scene resident, as in the server)
    plain = {}
    for scene, (dv, dims, level, _, isos) in scenes.items():
        for iso in isos:
            c = vt.Context(0)
            n, sizes = c.octree_build(dv, dims, iso, False, level)
            codes = pf.candidate_codes(c.octree_runs(level, sizes[level]))
            c.cache_create(level, 32, 3, max(1, int(codes.size)), 0)
            c.cache_attach_device_volume(dv, dims)
            c.cache_load_bricks(codes)
            c.set_march_options(skip_empty=True, max_defer=args.max_defer)
            plain[(scene, iso)] = (c, vt.render_params(level, level, iso))
    p_rgba, p_hits = ctx.alloc(1920 * 1080 * 16), ctx.alloc(1920 * 1080 * 32)

    def plain_replay(reqs):
        """Requests in arrival order, each rendered alone by vt_render_frame
        once it has arrived; latency from arrival to completion."""
        lat, busy = [], 0.0
        t0 = time.perf_counter()
        for q in reqs:
            wait = q["arrival_s"] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            c, prm = plain[(q["scene"], q["iso"])]
            f0 = time.perf_counter()
            c.render_frame(q["cam"], prm, p_rgba, p_hits)
            c.synchronize()
            f1 = time.perf_counter()
            busy += f1 - f0
            lat.append((f1 - t0) * 1e3 - q["arrival_s"] * 1e3)
        wall = time.perf_counter() - t0
        rays = sum(q["cam"].W * q["cam"].H for q in reqs)
        return np.array(lat), rays, wall, busy

    out = {"scenes": {"fbm_1024": {"resident_bricks": nA, "isos": [-0.2, 0.0, 0.2]},
                      "sdf_2048": {"resident_bricks": nB, "isos": [0.0]}}}
    for rate in args.rates:
        rng = np.random.default_rng(int(rate))
        n = max(16, int(rate * args.serve_seconds))
        walks = {}
        reqs = []
        t = 0.0
        for k in range(n):
            t += rng.exponential(1.0 / rate)
            viewer = int(rng.integers(32))
            W, H = res_set[int(rng.integers(4))]
            scene = sA if viewer % 2 == 0 else sB
            _, dims, _, radius, isos = scenes[scene]
            step = walks.get(viewer, 0)
            walks[viewer] = step + 1
            cam = random_walk(W, H, dims[0], step + 1, seed=7000 + viewer, radius=radius)[-1]
            iso = float(rng.choice(isos))
            reqs.append({"scene": scene, "cam": cam, "iso": iso, "arrival_s": t})
        # the first request of each (scene, iso) keeps its outputs: the plain
        # renderer must give the same bytes
        keep = {}
        for k, q in enumerate(reqs):
            keep.setdefault((q["scene"], q["iso"]), k)
        for k in keep.values():
            reqs[k]["keep"] = True
        row = {"requests": n}
        for mode, name in ((0, "batched"), (1, "per_request")):
            results, st, outs = sv.run(reqs, mode=mode)
            lat = np.array([r["latency_ms"] for r in results])
            row[name] = {"p50_ms": round(float(np.percentile(lat, 50)), 3),
                         "p99_ms": round(float(np.percentile(lat, 99)), 3),
                         "rays_per_s": round(st["rays"] / st["wall_s"], 1),
                         "busy_rays_per_s": round(st["rays"] / max(st["busy_s"], 1e-9), 1),
                         "launches": st["launches"], "wall_s": round(st["wall_s"], 3)}
            log(f"serving rate {rate}/s {name}: {row[name]}")
        lat, rays, wall, busy = plain_replay(reqs)
        row["plain_render_frame"] = {"p50_ms": round(float(np.percentile(lat, 50)), 3),
                                     "p99_ms": round(float(np.percentile(lat, 99)), 3),
                                     "rays_per_s": round(rays / wall, 1), "busy_rays_per_s": round(rays / busy, 1),
                                     "wall_s": round(wall, 3)}
        same = True
        for k in keep.values():
            q = reqs[k]
            c, prm = plain[(q["scene"], q["iso"])]
            R = q["cam"].W * q["cam"].H
            c.render_frame(q["cam"], prm, p_rgba, p_hits)
            c.synchronize()
            same = same and (p_hits.download(vt.HIT_DTYPE, (R,)).tobytes() == outs[k][1].tobytes() and
                             p_rgba.download(np.float32, (R, 4)).tobytes() == outs[k][0].tobytes())
        row["plain_render_frame"]["same_bytes_as_batched"] = same
        row["batched_over_plain"] = round(row["batched"]["rays_per_s"] / row["plain_render_frame"]["rays_per_s"], 3)
        log(f"serving rate {rate}/s plain_render_frame: {row['plain_render_frame']}")
        out[str(rate)] = row
    sv.close()
    for c, _ in plain.values():
        c.cache_destroy()
        c.close()
    for b in (p_rgba, p_hits, dA, dB):
        b.free()
    return out


def serving_leg(ctx, args, cams, prm, net, gpu_slots, out_bufs, R):
    """The prefetch+serving form: the path's frames submitted as requests to
    frames-in-flight serving (vt_frames) with the learned prefetcher loading
    the next waiting frame's bricks while the current launch runs.  A window
    of frames_in_flight + 2 frames is kept submitted (each renders into its
    own buffer); value = timed frames x rays / wall time of the timed
    frames, latency = submission to completion of each frame.  The warm-up
    frames are rendered one at a time by vt_render_frame before, so every
    k_render_multi / k_render_multi_wide launch of the run is a timed one."""
    J = args.frames_in_flight
    win = J + 2
    ring = [(ctx.alloc(R * 16), ctx.alloc(R * 32)) for _ in range(win)]
    nfr = len(cams)
    lat, passes = [], []
    for k in range(args.warmup):
        ctx.render_frame(cams[k], prm, *out_bufs.get(k, ring[k % win]))
    ctx.synchronize()
    with vt.Frames(ctx, max_frames=J, max_rays=R) as fr:
        fr.set_prefetch(net, gpu_slots, args.prefetch_min_prob)
        t0 = time.perf_counter()
        pending = []
        for k in range(args.warmup, nfr):
            if len(pending) == win:   # the oldest frame's buffer is reused next
                _, st = fr.poll(pending.pop(0))
                lat.append(st.ms_total)
                passes.append(st.passes)
            pending.append(fr.submit(cams[k], prm, *out_bufs.get(k, ring[k % win])))
        for t in pending:
            _, st = fr.poll(t)
            lat.append(st.ms_total)
            passes.append(st.passes)
        wall = time.perf_counter() - t0
        agg = fr.stats()
    for r_, h_ in ring:
        r_.free()
        h_.free()
    n = nfr - args.warmup
    m = np.array(lat)
    return {"value": round(R * n / wall, 1), "p50_ms": round(float(np.percentile(m, 50)), 3),
            "p99_ms": round(float(np.percentile(m, 99)), 3),
            "ms_per_frame": round(wall * 1e3 / n, 4),
            "demand_misses_per_frame": (agg["bricks_uploaded"]) / n,
            "early_loads_per_frame": agg["bricks_early"] / n,
            "prefetched_per_frame": agg["bricks_prefetched"] / n,
            "launches_per_frame": agg["launches"] / n, "launches_per_frame_in_flight": float(np.mean(passes)),
            "kernel_ms_per_frame": round(agg["ms_kernel"] / n, 4),
            "frames_in_flight": J, "submit_window": win, "latency": "submission to completion",
            "_kernel_ms": agg["ms_kernel"], "_launches": agg["launches"]}


def run_c3(ctx, args):
    """Config 3, the headline: 2048^3 signed-distance field (seeded union of
    200 spheres and tori, radii U[0.01, 0.08] of the extent, plus
    0.02-amplitude fBm), iso 0, 32^3 bricks (levels 0..6, 262144 bricks),
    HDF5 in 32^3 chunks; device cache 8 GB, pinned host tier 32 GB; a seeded
    random-walk camera path at 1920x1080.  The prefetcher is trained on the
    visited-brick traces of other seeded paths (rendered at 480x270 from the
    volume in HBM before the file is written).  The path is then rendered,
    each leg from a cold device cache: plain (one frame at a time), prefetch
    (a prefetch before every frame), and the rest as FORMS lists.  Frames of
    every leg are compared with the plain leg's bytes on sampled frames."""
    from visualtur_amd import prefetch as pf
    from visualtur_amd.volumes import C3_SEED, sdf_shapes

    D, level_c, inc, iso = 2048, 6, 3, 0.0
    dims = (D, D, D)
    radius = (1.3, 2.2)
    Sb = brick_bytes(32, inc)
    t0 = time.perf_counter()
    d_vol = ctx.alloc(4 * D ** 3)
    ctx.volume_generate(vt.GEN_SDF, dims, C3_SEED, d_vol, shapes=sdf_shapes(200, C3_SEED, D), n=D, octaves=5,
                        base_freq=1.0 / 32.0, gain=0.5, lacunarity=2.0, scale=0.02)
    log(f"c3 volume generated in {time.perf_counter() - t0:.2f}s")
    t0 = time.perf_counter()
    n, sizes = ctx.octree_build(d_vol, dims, iso, False, level_c)
    runs = [ctx.octree_runs(l, sizes[l]) for l in range(level_c + 1)]
    cand = pf.candidate_codes(runs[level_c])
    log(f"c3 octree built in {time.perf_counter() - t0:.2f}s (n={n}, bricks with nodes {cand.size})")
    prm = vt.render_params(level_c, level_c, iso)
    # training traces from the volume in HBM (every brick with nodes fits)
    t0 = time.perf_counter()
    ctx.cache_create(level_c, 32, inc, int(cand.size), 0)
    ctx.cache_attach_device_volume(d_vol, dims)
    tcams = [c for s in range(args.c3_train_paths)
             for c in pf.random_walk(480, 270, D, args.c3_train_frames, seed=30000 + s, radius=radius)]
    traces = pf.collect_traces(ctx, tcams, prm, level_c)
    net = vt.MLP(ctx, seed=13, max_batch=65536)
    X, Y = pf.build_dataset(ctx, net, tcams, traces, cand, n, level_c, neg_per_pos=3, seed=3, max_pos=1024)
    losses, _ = pf.train(ctx, net, X, Y, steps=args.c3_train_steps, batch=8192, lr=1e-3)
    train_s = time.perf_counter() - t0
    log(f"c3 prefetcher trained in {train_s:.1f}s: {len(tcams)} frames of {args.c3_train_paths} paths, "
        f"{X.shape[0]} samples, loss {losses[:10].mean():.3f} -> {losses[-10:].mean():.3f}")
    del X, Y
    tmpdir = tempfile.mkdtemp(prefix="vt_c3_")
    h5 = os.path.join(tmpdir, "c3.h5")
    t0 = time.perf_counter()
    ctx.volume_write_hdf5(d_vol, dims, h5, "value", 32)
    log(f"c3 hdf5 written in {time.perf_counter() - t0:.2f}s")
    vol_host = None if args.no_cpu else d_vol.download(np.float32, dims)
    ctx.cache_destroy()
    d_vol.free()
    gpu_slots, cpu_slots = int(8e9 // Sb), int(32e9 // Sb)
    W, H = 1920, 1080
    R = W * H
    nfr = args.warmup + args.steps
    cams = pf.random_walk(W, H, D, nfr, seed=C3_SEED, radius=radius)
    timed = list(range(args.warmup, nfr))
    sampled = sorted(set(timed[::8] + [timed[-1]]))   # frames whose bytes every leg must reproduce
    out_bufs = {k: (ctx.alloc(R * 16), ctx.alloc(R * 32)) for k in sampled}
    d_rgba, d_hits = ctx.alloc(R * 16), ctx.alloc(R * 32)

    def cache_setup():
        ctx.cache_create(level_c, 32, inc, gpu_slots, cpu_slots)
        ctx.cache_attach_hdf5(h5, "value")
        march_options(ctx, args)
        if args.wide_lanes is not None or args.wide_first is not None:
            ctx.set_wide_march(*wide_budgets(args))
        if args.early_reserve is not None:
            ctx.set_early_loads(args.early_reserve)
        if args.first_nodes is not None:
            ctx.set_first_pass_budget(args.first_nodes)
        if args.supertile is not None:
            ctx.set_first_pass_tiles(args.supertile)
        t = time.perf_counter()
        ctx.cache_prefetch_host(wait=True)
        return time.perf_counter() - t

    legs, frames = {}, {}
    for form in ("plain", "prefetch", "serving"):
        fill_s = cache_setup()
        log(f"c3 {form}: host tier filled from hdf5 in {fill_s:.2f}s, {ctx.cache_stats()['cpu_resident']} bricks")
        if form == "serving":
            legs[form] = serving_leg(ctx, args, cams, prm, net, gpu_slots, out_bufs, R)
        else:
            use = net if form == "prefetch" else None
            ms, pre_ms, uploads, passes, pre_n, kms = [], [], [], [], [], []
            for k in range(nfr):
                if k == args.warmup:
                    ctx.synchronize()
                    t_start = time.perf_counter()
                rg, hi = out_bufs.get(k, (d_rgba, d_hits))
                f0 = time.perf_counter()
                pre = use.prefetch(cams[k], gpu_slots, args.prefetch_min_prob) if use is not None else 0
                f1 = time.perf_counter()
                st = ctx.render_frame(cams[k], prm, rg, hi)
                f2 = time.perf_counter()
                if k >= args.warmup:
                    ms.append((f2 - f0) * 1e3)
                    pre_ms.append((f1 - f0) * 1e3)
                    uploads.append(st.bricks_uploaded)
                    passes.append(st.passes)
                    kms.append(st.ms_kernel)
                    pre_n.append(pre)
            ctx.synchronize()
            wall = time.perf_counter() - t_start
            m = np.array(ms)
            legs[form] = {"value": round(R * len(ms) / wall, 1), "p50_ms": round(float(np.percentile(m, 50)), 3),
                          "p99_ms": round(float(np.percentile(m, 99)), 3),
                          "demand_misses_per_frame": float(np.mean(uploads)), "passes_per_frame": float(np.mean(passes)),
                          "prefetched_per_frame": float(np.mean(pre_n)),
                          "prefetch_ms_per_frame": round(float(np.mean(pre_ms)), 3),
                          "kernel_ms_per_frame": round(float(np.mean(kms)), 4),
                          "_kernel_ms": float(np.sum(kms)), "_launches": int(np.sum(passes))}
        log(f"c3 {form}: { {k: v for k, v in legs[form].items() if not k.startswith('_')} }")
        frames[form] = {k: (out_bufs[k][1].download(vt.HIT_DTYPE, (R,)), out_bufs[k][0].download(np.float32, (R, 4)))
                        for k in sampled}
        ctx.cache_destroy()
    # bricks the reference algorithm marches in the timed frames (untimed replay, no deferral)
    cache_setup()
    U, samples = trace_bricks(ctx, args, [cams[k] for k in timed], prm, level_c, d_rgba, d_hits, collect_stats=True)
    ctx.cache_destroy()
    for form in legs:
        legs[form]["device_miss_rate"] = round(legs[form]["demand_misses_per_frame"] / max(1.0, float(np.mean(U))), 5)
    same = {form: all(frames[form][k][0].tobytes() == frames["plain"][k][0].tobytes() and
                      frames[form][k][1].tobytes() == frames["plain"][k][1].tobytes() for k in sampled)
            for form in legs}
    head = "serving"
    octree_bytes = 8 * sum(r.size for r in runs)
    B = sum(u * Sb + R * 48 + octree_bytes for u in U)
    achieved = B / (legs[head]["_kernel_ms"] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic("c3"),
                "kernel": "k_render_multi+k_render_multi_wide" if head == "serving" else RENDER_KERNELS, "leg": head,
                "samples_per_frame": int(np.mean(samples)),
                "avg_launch_ms": round(legs[head]["_kernel_ms"] / legs[head]["_launches"], 4),
                "bytes_per_launch": int(B / legs[head]["_launches"]),
                "bricks_marched_per_frame": int(np.mean(U)), "brick_bytes": Sb}
    for form in legs:
        for key in [k for k in legs[form] if k.startswith("_")]:
            del legs[form][key]
    out = {"metric": METRIC, "value": legs[head]["value"], "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(float(R / legs[head]["value"] * 1e3), 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "2048^3 SDF (200 spheres/tori + 0.02 fBm), iso 0, HDF5 32^3 chunks, 32^3 bricks + "
                                  "3-voxel halo, device cache 8 GB, pinned host tier 32 GB, 1920x1080 random walk",
                      "form": f"{head}: prefetch + frames in flight (plain and prefetch-only beside)",
                      "frames_in_flight": args.frames_in_flight, "bricks_with_nodes": int(cand.size),
                      "gpu_slots": gpu_slots, "cpu_slots": cpu_slots},
           "p50_ms": legs[head]["p50_ms"], "p99_ms": legs[head]["p99_ms"],
           "device_miss_rate": legs[head]["device_miss_rate"],
           "legs": legs, "speedup": {f: round(legs[f]["value"] / legs["plain"]["value"], 3) for f in legs},
           "identical_sampled_frames": same, "sampled_frames": len(sampled),
           "bricks_marched_per_frame": float(np.mean(U)), "roofline": roofline,
           "train": {"paths": args.c3_train_paths, "frames": len(tcams), "seconds": round(train_s, 2),
                     "final_loss": round(float(losses[-10:].mean()), 4)}, "cpu_baseline": None}
    if vol_host is not None:
        from oracle import vto

        w = Workload()
        w.prm = prm
        objs = (vto.Volume(vol_host), vto.Octree(runs, n, level_c), vto.params(level_c, level_c, iso))
        cache_setup()
        out["parity"] = parity_sample(w, ctx, cams[-1], objs, args.parity_rays, seed=3)
        ctx.cache_destroy()
        # the reference algorithm on this host's cores: whole frames of the
        # timed path, in order, for about cpu_seconds (all bricks in memory)
        cpu = host_cpu()
        t0 = time.perf_counter()
        done = 0
        for k in timed:
            vto.render(*objs[:2], cams[k], objs[2], nthreads=cpu["usable"])
            done += 1
            if time.perf_counter() - t0 >= args.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(done * R / dt, 1), "unit": "rays/s", "cores": cpu["usable"],
                               "kind": "port", "host": cpu["model"], "nproc": cpu["nproc"],
                               "sample": f"the first {done} timed frames of the path (1920x1080 each), {dt:.1f}s "
                                         f"wall, OpenMP over 64-ray chunks, all bricks in memory"}
    for b in [d_rgba, d_hits] + [x for p in out_bufs.values() for x in p]:
        b.free()
    net.close()
    shutil.rmtree(tmpdir, ignore_errors=True)
    return out


def morton(level: int, bx, by, bz) -> np.ndarray:
    """Codes of level-`level` nodes (leading 1 bit, then x, y, z bits interleaved, x highest)."""
    bx, by, bz = (np.asarray(c, np.uint64) for c in (bx, by, bz))
    code = np.full(np.broadcast(bx, by, bz).shape, np.uint64(1) << np.uint64(3 * level), np.uint64)
    for i in range(level):
        b = np.uint64(i)
        code |= (((bx >> b) & np.uint64(1)) << np.uint64(3 * i + 2)) | \
                (((by >> b) & np.uint64(1)) << np.uint64(3 * i + 1)) | (((bz >> b) & np.uint64(1)) << np.uint64(3 * i))
    return code


def in_runs(runs: np.ndarray, codes: np.ndarray) -> np.ndarray:
    """Membership of `codes` in a run list (sorted inclusive [start, end] pairs)."""
    starts, ends = runs[0::2], runs[1::2]
    i = np.searchsorted(starts, codes, side="right").astype(np.int64) - 1
    return (i >= 0) & (codes <= ends[np.maximum(i, 0)])


def fly_through(W: int, H: int, frames: int, X: int, Y: int, Z: int, dy: float = 0.0, dz: float = 0.0,
                phase: float = 0.0):
    """A fly-through: enters the volume through the x = 0 face at mid
    height and depth (offset by dy, dz of the extents) and flies along +x to
    85 % of its length with a slow sideways and vertical weave (shifted by
    `phase` of a period), looking along the path."""
    def pos(t):
        return np.array([-0.15 * X + 1.0 * X * t,
                         (0.5 + dy) * Y + 0.15 * Y * math.sin(2 * math.pi * (t + phase)),
                         (0.5 + dz) * Z + 0.1 * Z * math.sin(4 * math.pi * (t + phase))], np.float64)
    cams = []
    for k in range(frames):
        t = k / max(1, frames - 1)
        eye = pos(t)
        look, up, right = vt.look_at(eye, pos(t + 0.02))
        cams.append(vt.make_camera(W, H, eye, look, up, right, fov_v_deg=45.0))
    return cams


def c5_viewer_server(ctx, args, d_vol, dims, level_c, iso):
    """The concurrent 1024^2 viewers of the large configuration: the volume
    becomes a serving scene for its one isovalue (only the bricks that hold
    nodes at it are resident), and seeded random-walk viewers issue requests
    as a Poisson stream.  Returns the server, the requests and the scene's
    size, after an untimed warm-up that builds the isovalue's tables and the
    captured launches."""
    from visualtur_amd import prefetch as pf
    from visualtur_amd.prefetch import random_walk

    sv = vt.Server(ctx, hbm_budget=2 << 30, max_tiles=65536)
    t0 = time.perf_counter()
    scene, resident = sv.add_scene_range(d_vol, dims, level_c, level_c, iso, iso)
    log(f"c5 viewers: scene of {resident} resident bricks in {time.perf_counter() - t0:.1f}s")
    rng = np.random.default_rng(0xC5)
    nv, per = args.c5_viewers, args.c5_viewer_frames
    walks = {v: random_walk(1024, 1024, dims[0], per, seed=9000 + v) for v in range(nv)}
    reqs, t = [], 0.0
    for k in range(nv * per):
        t += rng.exponential(1.0 / args.c5_viewer_rate)
        v = k % nv
        reqs.append({"scene": scene, "cam": walks[v][k // nv], "iso": iso, "arrival_s": t, "keep": k < 8})
    t0 = time.perf_counter()
    for mode in (0, 1):
        sv.run([dict(q, arrival_s=0.0, keep=False) for q in reqs[:nv]], mode=mode)
    log(f"c5 viewers: warm-up in {time.perf_counter() - t0:.2f}s")
    return sv, reqs, {"viewers": nv, "requests": len(reqs), "rate_per_s": args.c5_viewer_rate,
                      "resident_bricks": resident, "resolution": "1024x1024"}


def viewer_row(res, st):
    lat = np.array([r["latency_ms"] for r in res])
    return {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "rays_per_s": round(st["rays"] / st["wall_s"], 1),
            "busy_rays_per_s": round(st["rays"] / max(st["busy_s"], 1e-9), 1), "launches": st["launches"]}


def run_c5(ctx, args):
    """Config 5 with the volume resident in HBM as the brick source:
    4096x4096x2048 float32 (128 GiB), a seeded turbulence-like field (64
    random-phase Fourier modes, |k|^-5/3 amplitudes), iso at the 70th
    percentile of a seeded 1% sample (41 random x-slabs); logical cube 4096^3
    with zero padding above z = 2048; 32^3 bricks (levels 0..7, 2^21 bricks)
    with a 3-voxel halo; device brick cache 48 GB; a 3840x2160 fly-through.
    The 128 GB HDF5 file is replaced by the volume in HBM (no disk here
    holds it): the pinned host tier holds every brick with nodes, filled once
    from it, so device misses cross PCIe as in the configuration.  The 16
    concurrent 1024^2 viewers are served from a scene holding the bricks with
    nodes at the isovalue, alone (batched and one launch per request) and
    beside the fly-through's frames in flight.
    Size-independent checks at full size: the generator against its float32
    formula and its own slabs, the builder's level-7 nodes against the cell
    test on one layer of bricks, hits unchanged with an 8 GB cache, and
    (unless --no-cpu) sampled rays against the oracle."""
    from visualtur_amd.volumes import C5_SEED, fourier_field, fourier_modes

    X, Y, Z = 4096, 4096, 2048
    level_c, inc = 7, 3
    Sb = brick_bytes(32, inc)
    modes = fourier_modes(64, C5_SEED, X)
    slab_n = Y * Z

    def gen(d, x0, nx):
        ctx.volume_generate(vt.GEN_FOURIER, (X, Y, Z), C5_SEED, d, shapes=modes, x0=x0, nx=nx)

    checks = {}
    # iso: 70th percentile of a seeded 1% voxel sample (41 of 4096 x-slabs)
    t0 = time.perf_counter()
    rng = np.random.default_rng(C5_SEED)
    xs = np.sort(rng.choice(X, 41, replace=False))
    slab = ctx.alloc(slab_n * 4)
    sample = np.empty((xs.size, slab_n), np.float32)
    for i, x in enumerate(xs):
        gen(slab, int(x), 1)
        sample[i] = slab.download(np.float32, (slab_n,))
    iso = float(np.float32(np.percentile(sample, 70)))
    log(f"c5 iso {iso:.6f} (70th percentile of {sample.size} voxels) in {time.perf_counter() - t0:.1f}s")
    # the volume
    t0 = time.perf_counter()
    d_vol = ctx.alloc(X * Y * Z * 4)
    gen(d_vol, 0, X)
    log(f"c5 volume generated in {time.perf_counter() - t0:.2f}s")
    # generator: the full buffer holds the slabs generated alone (64-bit
    # offsets), and they equal the float32 formula
    same = all(np.array_equal(d_vol.download_at(int(x) * slab_n * 4, np.float32, (slab_n,)).view(np.uint32),
                              sample[i].view(np.uint32)) for i, x in ((0, xs[0]), (xs.size - 1, xs[-1])))
    yz = rng.integers(0, [Y, Z], (4096, 2))
    ref = fourier_field(modes, np.full(4096, xs[-1]), yz[:, 0], yz[:, 1])
    err = float(np.max(np.abs(sample[-1].reshape(Y, Z)[yz[:, 0], yz[:, 1]] - ref)))
    checks["generator"] = {"slabs_equal": bool(same), "max_abs_err_vs_formula": err, "points": 4096}
    del sample
    # octree on the GPU
    t0 = time.perf_counter()
    n, sizes = ctx.octree_build(d_vol, (X, Y, Z), iso, False, level_c)
    runs = [ctx.octree_runs(l, sizes[l]) for l in range(level_c + 1)]
    log(f"c5 octree built in {time.perf_counter() - t0:.2f}s (n={n}, level-{level_c} runs {sizes[level_c] // 2}, "
        f"leaf runs {sizes[-1] // 2})")
    # builder: every level-7 node of the brick layer x in [2048, 2080)
    # against the cell test (corners on both sides of iso, zero outside)
    t0 = time.perf_counter()
    bx, x0 = 64, 2048
    layer = ctx.alloc(33 * slab_n * 4)
    gen(layer, x0, 33)
    below = np.full((33, Y + 1, Z + 1), np.float32(0.0) < np.float32(iso))
    below[:, :Y, :Z] = layer.download(np.float32, (33, Y, Z)) < np.float32(iso)
    layer.free()
    any_b = np.zeros((32, Y, Z), bool)
    all_b = np.ones((32, Y, Z), bool)
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                v = below[dx:dx + 32, dy:dy + Y, dz:dz + Z]
                any_b |= v
                all_b &= v
    cell_on = any_b & ~all_b
    del below, any_b, all_b
    expect = np.zeros((128, 128), bool)   # (by, bz); bz >= 64 lies in the zero padding
    expect[:, :64] = cell_on.reshape(32, 128, 32, 64, 32).any(axis=(0, 2, 4))
    del cell_on
    by, bz = np.meshgrid(np.arange(128), np.arange(128), indexing="ij")
    got = in_runs(runs[level_c], morton(level_c, np.full(by.shape, bx), by, bz))
    checks["builder_layer"] = {"bricks": int(expect.size), "with_nodes": int(expect.sum()),
                               "mismatches": int((got != expect).sum())}
    log(f"c5 builder check {checks['builder_layer']} in {time.perf_counter() - t0:.1f}s")
    prm = vt.render_params(level_c, level_c, iso)
    W, H = 3840, 2160
    R = W * H
    nfr = args.warmup + args.steps
    cams = fly_through(W, H, nfr, X, Y, Z)
    timed = list(range(args.warmup, nfr))
    sampled = sorted(set(timed[::16] + [timed[-1]]))
    out_bufs = {k: (ctx.alloc(R * 16), ctx.alloc(R * 32)) for k in sampled}
    d_rgba, d_hits = ctx.alloc(R * 16), ctx.alloc(R * 32)
    gpu_slots, cpu_slots = int(48e9 // Sb), int(96e9 // Sb)
    with_nodes = int(sum(int(runs[level_c][j + 1] - runs[level_c][j]) + 1 for j in range(0, runs[level_c].size, 2)))

    def cache_setup(slots):
        # brick source: a pinned host tier holding every brick with nodes,
        # filled once from the volume in HBM, so device misses cross PCIe
        t = time.perf_counter()
        ctx.cache_create(level_c, 32, inc, slots, cpu_slots)
        ctx.cache_attach_host_from_device(d_vol, (X, Y, Z))
        march_options(ctx, args)
        if args.supertile is not None:
            ctx.set_first_pass_tiles(args.supertile)
        return time.perf_counter() - t

    # the learned prefetcher, trained on other fly-throughs (other heights,
    # depths and weave phases) rendered at 960x540 from the volume in HBM
    net, train = None, None
    if args.prefetch:
        from visualtur_amd import prefetch as pf

        t0 = time.perf_counter()
        cand = pf.candidate_codes(runs[level_c])
        ctx.cache_create(level_c, 32, inc, int(cand.size), 0)
        ctx.cache_attach_device_volume(d_vol, (X, Y, Z))
        tcams = [c for k, (dy, dz, ph) in enumerate([(-0.2, -0.15, 0.1), (0.2, 0.1, 0.3), (-0.1, 0.2, 0.55),
                                                     (0.15, -0.2, 0.75), (0.05, 0.05, 0.9), (-0.25, 0.0, 0.4)])
                 for c in fly_through(960, 540, 32, X, Y, Z, dy, dz, ph)]
        traces = pf.collect_traces(ctx, tcams, prm, level_c)
        net = vt.MLP(ctx, seed=17, max_batch=65536)
        Xf, Yf = pf.build_dataset(ctx, net, tcams, traces, cand, n, level_c, neg_per_pos=3, seed=5, max_pos=4096)
        losses, _ = pf.train(ctx, net, Xf, Yf, steps=args.c3_train_steps, batch=8192, lr=1e-3)
        ctx.cache_destroy()
        train = {"frames": len(tcams), "samples": int(Xf.shape[0]), "seconds": round(time.perf_counter() - t0, 2),
                 "final_loss": round(float(losses[-10:].mean()), 4)}
        log(f"c5 prefetcher trained: {train}")
        del Xf, Yf
    # plain: one frame at a time
    fill_s = cache_setup(gpu_slots)
    log(f"c5 plain: pinned host tier of {ctx.cache_stats()['cpu_resident']} bricks filled from HBM in {fill_s:.1f}s")
    for k in range(args.warmup):
        ctx.render_frame(cams[k], prm, *out_bufs.get(k, (d_rgba, d_hits)))
    ctx.synchronize()
    ms, kms, passes, uploads = [], [], [], []
    t0 = time.perf_counter()
    for k in timed:
        f0 = time.perf_counter()
        st = ctx.render_frame(cams[k], prm, *out_bufs.get(k, (d_rgba, d_hits)))
        ms.append((time.perf_counter() - f0) * 1e3)
        kms.append(st.ms_kernel)
        passes.append(st.passes)
        uploads.append(st.bricks_uploaded)
    ctx.synchronize()
    total_s = time.perf_counter() - t0
    frames = {"plain": {k: (out_bufs[k][1].download(vt.HIT_DTYPE, (R,)), out_bufs[k][0].download(np.float32, (R, 4)))
                        for k in sampled}}
    last = frames["plain"][timed[-1]]
    resident = ctx.cache_stats()["gpu_resident"]
    # bricks the reference algorithm marches in the timed frames (trace replay, untimed)
    U = trace_bricks(ctx, args, [cams[k] for k in timed], prm, level_c, d_rgba, d_hits)
    # the last frame again through an 8 GB cache, from cold: the same bytes
    ctx.cache_destroy()
    cache_setup(int(8e9 // Sb))
    st = ctx.render_frame(cams[-1], prm, d_rgba, d_hits)
    small = (d_hits.download(vt.HIT_DTYPE, (R,)), d_rgba.download(np.float32, (R, 4)))
    checks["small_cache_same_frame"] = {"identical": bool(small[0].tobytes() == last[0].tobytes() and
                                                           small[1].tobytes() == last[1].tobytes()),
                                        "hit_fraction": round(float((last[0]["node"] != 0).mean()), 4),
                                        "passes": int(st.passes), "uploads": int(st.bricks_uploaded)}
    ctx.cache_destroy()
    plain = {"value": round(R * len(timed) / total_s, 1), "p50_ms": round(float(np.percentile(ms, 50)), 3),
             "p99_ms": round(float(np.percentile(ms, 99)), 3), "kernel_ms_per_frame": round(float(np.mean(kms)), 4),
             "passes_per_frame": float(np.mean(passes)), "bricks_uploaded_per_frame": float(np.mean(uploads)),
             "device_miss_rate": round(float(sum(uploads)) / max(1, sum(U)), 5), "resident_after": int(resident)}
    # the viewers alone (batched and one launch per request), then the
    # serving form: the fly-through as frames in flight with the prefetcher
    # and the viewers' request stream served beside it, both on the GPU at once
    viewers, serving = None, None
    sv = vreqs = None
    if args.c5_viewers > 0:
        sv, vreqs, viewers = c5_viewer_server(ctx, args, d_vol, (X, Y, Z), level_c, iso)
        kept = {}
        for mode, name in ((0, "batched_alone"), (1, "per_request_alone")):
            res, vst, o = sv.run(vreqs, mode=mode)
            kept[name] = o
            viewers[name] = viewer_row(res, vst)
            log(f"c5 viewers {name}: {viewers[name]}")
        viewers["same_bytes_batched_and_alone"] = all(
            kept["batched_alone"][k][1].tobytes() == kept["per_request_alone"][k][1].tobytes() and
            kept["batched_alone"][k][0].tobytes() == kept["per_request_alone"][k][0].tobytes()
            for k in kept["batched_alone"])
    if args.c5_frames_in_flight > 0:
        import threading

        cache_setup(gpu_slots)
        J = args.c5_frames_in_flight
        win = J + 2
        ring = [(ctx.alloc(R * 16), ctx.alloc(R * 32)) for _ in range(win)]
        beside = {}
        th = threading.Thread(target=lambda: beside.update(zip(("res", "st", "out"), sv.run(vreqs, mode=0))))
        # warm-up frames one at a time, so every multi-frame launch is timed
        for k in range(args.warmup):
            ctx.render_frame(cams[k], prm, *out_bufs.get(k, ring[k % win]))
        ctx.synchronize()
        with vt.Frames(ctx, max_frames=J, max_rays=R) as fr:
            if net is not None:
                fr.set_prefetch(net, gpu_slots, args.prefetch_min_prob)
            if sv is not None:
                th.start()
            t0 = time.perf_counter()
            pending, lat = [], []
            for k in timed:
                if len(pending) == win:
                    lat.append(fr.poll(pending.pop(0))[1].ms_total)
                pending.append(fr.submit(cams[k], prm, *out_bufs.get(k, ring[k % win])))
            for t in pending:
                lat.append(fr.poll(t)[1].ms_total)
            wall = time.perf_counter() - t0
            agg = fr.stats()
        if sv is not None:
            th.join()
            viewers["batched_beside_flythrough"] = viewer_row(beside["res"], beside["st"])
            viewers["beside_same_bytes"] = all(
                beside["out"][k][1].tobytes() == kept["per_request_alone"][k][1].tobytes() for k in beside["out"])
            log(f"c5 viewers beside the fly-through: {viewers['batched_beside_flythrough']}")
        for r_, h_ in ring:
            r_.free()
            h_.free()
        n_t = len(timed)
        serving = {"value": round(R * n_t / wall, 1), "p50_ms": round(float(np.percentile(lat, 50)), 3),
                   "p99_ms": round(float(np.percentile(lat, 99)), 3), "ms_per_frame": round(wall * 1e3 / n_t, 4),
                   "kernel_ms_per_frame": round(agg["ms_kernel"] / n_t, 4),
                   "launches_per_frame": agg["launches"] / n_t,
                   "bricks_uploaded_per_frame": agg["bricks_uploaded"] / n_t,
                   "prefetched_per_frame": agg["bricks_prefetched"] / n_t, "frames_in_flight": J,
                   "viewers_beside": sv is not None, "latency": "submission to completion (window of J + 2 frames)",
                   "identical_sampled_frames": all(
                       out_bufs[k][1].download(vt.HIT_DTYPE, (R,)).tobytes() == frames["plain"][k][0].tobytes() and
                       out_bufs[k][0].download(np.float32, (R, 4)).tobytes() == frames["plain"][k][1].tobytes()
                       for k in sampled),
                   "_kernel_ms": agg["ms_kernel"], "_launches": agg["launches"]}
        serving["speedup"] = round(serving["value"] / plain["value"], 3)
        log(f"c5 serving: { {k: v for k, v in serving.items() if not k.startswith('_')} }")
        ctx.cache_destroy()
    if sv is not None:
        sv.close()
    if net is not None:
        net.close()
    head = serving if serving is not None else plain
    octree_bytes = 8 * sum(r.size for r in runs)
    B = sum(u * Sb + R * 48 + octree_bytes for u in U)
    h_kms = head["_kernel_ms"] if serving is not None else sum(kms)
    h_launches = head["_launches"] if serving is not None else sum(passes)
    achieved = B / (h_kms * 1e-3) / 1e9
    if serving is not None:
        del serving["_kernel_ms"], serving["_launches"]
    out = {"metric": METRIC, "value": head["value"], "unit": "rays/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(R / head["value"] * 1e3, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "4096x4096x2048 turbulence (64 Fourier modes, k^-5/3), iso = 70th percentile, "
                                  "32^3 bricks + 3-voxel halo, device cache 48 GB, pinned host tier (the bricks with "
                                  "nodes, filled from the volume in HBM), 3840x2160 fly-through, 16 1024^2 viewers",
                      "form": "serving: prefetch + frames in flight, viewers beside" if serving else "plain",
                      "iso": iso, "gpu_slots": gpu_slots, "cpu_slots": cpu_slots, "bricks": 1 << (3 * level_c),
                      "bricks_with_nodes": with_nodes},
           "p50_ms": head["p50_ms"], "p99_ms": head["p99_ms"], "plain": plain, "serving": serving,
           "bricks_marched_per_frame": float(np.mean(U)),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic("c5"),
                        "kernel": "k_render_multi+k_render_multi_wide" if serving else RENDER_KERNELS,
                        "avg_launch_ms": round(h_kms / h_launches, 4), "bytes_per_launch": int(B / h_launches),
                        "bricks_marched_per_frame": int(np.mean(U)), "brick_bytes": Sb},
           "checks": checks, "viewers": viewers, "train": train, "cpu_baseline": None}
    if not args.no_cpu:
        from oracle import vto

        # the volume in host memory for the oracle (128 GiB, in 8 GiB pieces)
        t0 = time.perf_counter()
        vol_host = np.empty((X, Y, Z), np.float32)
        step = 256
        for xa in range(0, X, step):
            part = vol_host[xa:xa + step]
            vt._check(vt.lib().vt_memcpy_d2h(ctx.handle, part.ctypes.data, d_vol.ptr + xa * slab_n * 4,
                                             part.nbytes), ctx.handle)
        log(f"c5 volume copied to the host in {time.perf_counter() - t0:.1f}s")
        cache_setup(gpu_slots)
        w = Workload()
        w.prm = prm
        objs = (vto.Volume(vol_host), vto.Octree(runs, n, level_c), vto.params(level_c, level_c, iso))
        out["parity"] = parity_sample(w, ctx, cams[-1], objs, args.parity_rays, seed=5)
        log(f"c5 parity {out['parity']}")
        threads = len(os.sched_getaffinity(0))
        ids = np.arange(0, R, max(1, R // 65536), dtype=np.int64)
        t0 = time.perf_counter()
        done = 0
        while time.perf_counter() - t0 < args.cpu_seconds:
            vto.render_list(*objs[:2], cams[-1], objs[2], ids)
            done += ids.size
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": done / dt, "unit": "rays/s", "cores": threads, "kind": "port",
                               "sample": f"{done} rays ({ids.size} evenly spaced rays of the last frame, repeated), "
                                         f"{dt:.1f}s wall, all bricks in memory"}
        ctx.cache_destroy()
        del objs, vol_host
    for b in [d_rgba, d_hits] + [x for pr in out_bufs.values() for x in pr]:
        b.free()
    slab.free()
    d_vol.free()
    return out


def cpu_baseline(w, cams, seconds: float, threads: int | None = None):
    """Reference algorithm (the C oracle) on this host's cores (or `threads`
    of them), over whole frames of the same workload until `seconds` of wall
    time have elapsed."""
    from oracle import vto

    cpu = host_cpu()
    threads = cpu["usable"] if threads is None else threads
    V = vto.Volume(w.oracle_volume())
    n, runs = w.oracle_runs()
    O = vto.Octree(runs, n, w.level_o)
    prm = vto.params(w.level_o, w.level_c, w.iso)
    cached = getattr(w, "cpu_cache", None)
    src = vto.H5Source(cached[0], "value") if cached else None
    lru = vto.BrickLRU(w.level_c, w.cube, w.inc, cached[1]) if cached else None
    rays = 0
    frames = 0
    misses = passes = 0
    t0 = time.perf_counter()
    while True:
        cam = cams[frames % len(cams)]
        if cached:
            # the reference frame loop: passes of traversal, an LRU cache of
            # the GPU cache's capacity read from the same HDF5 file, march
            _, _, st = vto.render_cached(O, cam, prm, lru, src, nthreads=threads)
            misses += st["misses"]
            passes += st["passes"]
        else:
            vto.render(V, O, cam, prm, nthreads=threads)
        rays += cam.W * cam.H
        frames += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    if src is not None:
        src.close()
        lru.close()
    how = (f"the reference pass loop with an LRU of {cached[1]} bricks read from the HDF5 file, kept across "
           f"frames from cold ({misses / frames:.0f} reads and {passes / frames:.1f} passes per frame)"
           if cached else "all bricks in memory")
    return {"value": rays / dt, "unit": "rays/s", "cores": threads, "kind": "port", "host": cpu["model"],
            "nproc": cpu["nproc"], "cache": "lru" if cached else "all resident",
            "sample": f"{frames} full {cams[0].W}x{cams[0].H} frames of the timed camera path, "
                      f"{dt:.1f}s wall, OpenMP over rays, {how}"}, (V, O, prm)


def parity_sample(w, ctx, cam, oracle_objs, nrays: int, seed: int):
    """Render `cam` on the GPU and compare `nrays` random rays with the oracle."""
    from oracle import vto

    V, O, prm = oracle_objs
    R = cam.W * cam.H
    d_rgba, d_hits = ctx.alloc(R * 16), ctx.alloc(R * 32)
    ctx.render_frame(cam, w.prm, d_rgba, d_hits)
    rgba = d_rgba.download(np.float32, (R, 4))
    hits = d_hits.download(vt.HIT_DTYPE, (R,))
    d_rgba.free()
    d_hits.free()
    # three quarters of the sample among rays the GPU reports as hits (the
    # rest uniform), so that marched rays dominate the check
    rng = np.random.default_rng(seed)
    hit_ids = np.nonzero(hits["node"] != 0)[0]
    k_hit = min(hit_ids.size, (3 * nrays) // 4)
    ids = np.unique(np.concatenate([rng.choice(hit_ids, size=k_hit, replace=False) if k_hit else
                                    np.zeros(0, np.int64), rng.choice(R, size=min(nrays - k_hit, R), replace=False)]))
    rc, hc = vto.render_list(V, O, cam, prm, ids)
    hg, rg = hits[ids], rgba[ids]
    exact = (hg["node"] == hc["node"]) & (hg["brick"] == hc["brick"]) & np.all(hg["voxel"] == hc["voxel"], axis=1)
    dg, dc = hg["depth"], hc["depth"]
    # bytewise, NaN matching NaN (a 0/0 secant gives NaN on both sides, with
    # a sign bit that differs between the GPU and the host)
    nan_d = np.isnan(dg) & np.isnan(dc)
    nan_c = np.isnan(rg) & np.isnan(rc)
    depth_bits = nan_d | (dg.view(np.uint32) == dc.view(np.uint32))
    rgba_bits = np.all(nan_c | (rg.view(np.uint32) == rc.view(np.uint32)), axis=1)
    depth_tol = nan_d | (np.abs(dg.astype(np.float64) - dc.astype(np.float64)) <= 1e-4)
    rgba_tol = np.all(nan_c | (np.abs(rg.astype(np.float64) - rc) <= 1.0 / 255), axis=1)
    bad = int((~(exact & depth_tol & rgba_tol)).sum())
    tol_only = int((exact & depth_tol & rgba_tol & ~(depth_bits & rgba_bits)).sum())
    return {"rays_checked": int(ids.size), "mismatches": bad, "within_tolerance_not_bitwise": tol_only,
            "hit_fraction": float((hc["node"] != 0).mean())}


def wide_budgets(args) -> tuple[int, int]:
    """Lane budgets (later passes, first pass); unset ones keep the library defaults."""
    return (1 << 18 if args.wide_lanes is None else args.wide_lanes,
            1 << 21 if args.wide_first is None else args.wide_first)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--c3-train-paths", type=int, default=512)
    ap.add_argument("--c3-train-frames", type=int, default=4)
    ap.add_argument("--c3-train-steps", type=int, default=1500)
    ap.add_argument("--c5-viewers", type=int, default=16, help="c5: concurrent 1024^2 viewers served beside (0: none)")
    ap.add_argument("--c5-viewer-frames", type=int, default=8)
    ap.add_argument("--c5-viewer-rate", type=float, default=200.0)
    ap.add_argument("--rates", type=float, nargs="*", default=[50, 100, 200, 400, 800])
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--parity-rays", type=int, default=4096)
    ap.add_argument("--prefetch", type=int, default=1, help="c2: also time the learned-prefetch form")
    ap.add_argument("--persistent", type=int, default=0, help="persistent render pass with lane refill")
    ap.add_argument("--max-defer-late", type=int, default=2, help="max-defer of the passes after the first (-1: same)")
    ap.add_argument("--wide-lanes", type=int, default=None,
                    help="lane budget of the wide march in later passes (0: off; default: the library's)")
    ap.add_argument("--supertile", type=int, default=None,
                    help="first pass: S x S blocks of 8x8 tiles per XCD (1: single tiles; default: the library's)")
    ap.add_argument("--first-nodes", type=int, default=None,
                    help="walk nodes per ray in the first pass before it parks (0: no limit; default: the library's)")
    ap.add_argument("--early-reserve", type=int, default=None,
                    help="free slots for loads beside the first pass (0: off; default: the library's)")
    ap.add_argument("--wide-first", type=int, default=None,
                    help="lane budget of the wide march in the first pass (0: off; default: the library's)")
    ap.add_argument("--max-defer", type=int, default=1,
                    help="absent-brick nodes a ray may walk past per pass (0 = park at the first)")
    ap.add_argument("--prefetch-min-prob", type=float, default=0.8)
    ap.add_argument("--frames-in-flight", type=int, default=32, help="c3 serving leg: frames sharing launches")
    ap.add_argument("--c2-source", default="bricks", choices=["bricks", "volume"],
                    help="c2 host tier source: brick-major file or the 32^3-chunked volume file")
    ap.add_argument("--c5-frames-in-flight", type=int, default=4, help="c5 serving leg: frames sharing launches")
    ap.add_argument("--train-paths", type=int, default=24)
    ap.add_argument("--train-frames", type=int, default=6)
    ap.add_argument("--train-steps", type=int, default=600)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 256 MiB streaming copy kernel first (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the multi-GPU (image-space sharded) form is not built; "
              "run with --gpus 1", file=sys.stderr)
        return 2

    ctx = vt.Context(0)
    if args.copy_probe:
        nb = 256 << 20
        a_, b_ = ctx.alloc(nb), ctx.alloc(nb)
        for _ in range(3):
            ctx.copy_probe(a_, b_, nb)
        a_.free()
        b_.free()
    if args.config == "c4":
        sweep = serving_sweep(ctx, args)
        rows = [r for k, r in sweep.items() if k != "scenes"]
        best = max(rows, key=lambda r: r["batched"]["rays_per_s"])
        out = {"metric": METRIC, "value": best["batched"]["rays_per_s"], "unit": "rays/s", "n_gpus": 1,
               "steps": sum(r["requests"] for r in rows), "warmup": 0, "ms_per_step": None,
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
               "data": "synthetic",
               "config": {"workload": "serving: 1024^3 fBm (iso -0.2/0/0.2) and 2048^3 SDF (iso 0) resident, 32 "
                                      "random-walk viewers, Poisson arrivals, mixed resolutions, per-request iso",
                          "form": "batched (continuous batching) vs per-request (one request per launch) vs plain "
                                  "vt_render_frame per request"},
               "serving": sweep}
        ctx.close()
        print(json.dumps(out), flush=True)
        return 0
    if args.config in ("c3", "c5"):
        out = run_c3(ctx, args) if args.config == "c3" else run_c5(ctx, args)
        ctx.close()
        print(json.dumps(out), flush=True)
        return 0
    w = setup_c2(ctx, args) if args.config == "c2" else setup_c1(ctx, args)
    march_options(ctx, args)
    if args.wide_lanes is not None or args.wide_first is not None:
        ctx.set_wide_march(*wide_budgets(args))
    if args.early_reserve is not None:
        ctx.set_early_loads(args.early_reserve)
    if args.first_nodes is not None:
        ctx.set_first_pass_budget(args.first_nodes)
    if args.supertile is not None:
        ctx.set_first_pass_tiles(args.supertile)
    cams = [w.cameras(k) for k in range(args.warmup + args.steps)]
    R = w.W * w.H
    d_rgba, d_hits = ctx.alloc(R * 16), ctx.alloc(R * 32)

    for k in range(args.warmup):
        ctx.render_frame(cams[k], w.prm, d_rgba, d_hits)
    ctx.synchronize()

    frame_ms, kernel_ms, passes, uploads, parked, early = [], [], [], [], [], []
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        f0 = time.perf_counter()
        st = ctx.render_frame(cams[k], w.prm, d_rgba, d_hits)
        frame_ms.append((time.perf_counter() - f0) * 1e3)
        kernel_ms.append(st.ms_kernel)
        passes.append(st.passes)
        uploads.append(st.bricks_uploaded)
        parked.append(st.rays_parked)
        early.append(st.bricks_early)
    ctx.synchronize()
    total_s = time.perf_counter() - t0
    rays_per_s = R * args.steps / total_s

    # algorithmic bytes per frame: distinct bricks the reference algorithm
    # marches (trace replay without deferral, untimed) x brick payload + 48 B
    # per ray + octree levels 0..levelO
    U, samples = trace_bricks(ctx, args, cams[args.warmup: args.warmup + args.steps], w.prm, w.level_c, d_rgba,
                              d_hits, collect_stats=True)
    octree_bytes = 8 * sum(r.size for r in w.runs[: w.level_o + 1])
    Sb = brick_bytes(w.cube, w.inc)
    B = [u * Sb + R * 48 + octree_bytes for u in U]
    launches = sum(passes)
    achieved_gbs = sum(B) / (sum(kernel_ms) * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved_gbs / HBM_PEAK_GBS, 5), "traffic": measured_traffic(w.name),
                "kernel": RENDER_KERNELS, "avg_launch_ms": round(sum(kernel_ms) / launches, 4),
                "bytes_per_launch": int(sum(B) / launches),
                "bricks_marched_per_frame": int(np.mean(U)), "brick_bytes": Sb,
                "samples_per_frame": int(np.mean(samples))}

    ms = np.array(frame_ms)
    out = {
        "metric": METRIC, "value": round(rays_per_s, 1), "unit": "rays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(float(ms.mean()), 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.desc, "form": "plain", **w.extra},
        "p50_ms": round(float(np.percentile(ms, 50)), 4), "p99_ms": round(float(np.percentile(ms, 99)), 4),
        "kernel_ms_per_frame": round(float(np.mean(kernel_ms)), 4),
        "passes_per_frame": float(np.mean(passes)), "bricks_uploaded_per_frame": float(np.mean(uploads)),
        "bricks_early_per_frame": float(np.mean(early)),
        "rays_parked_per_frame": float(np.mean(parked)), "max_defer": args.max_defer,
        "persistent": args.persistent, "max_defer_late": args.max_defer_late, "wide_lanes": args.wide_lanes, "wide_first": args.wide_first, "early_reserve": args.early_reserve, "first_nodes": args.first_nodes,
        "bricks_marched_per_frame": float(np.mean(U)),
        "device_miss_rate": round(float(sum(uploads)) / max(1, sum(U)), 5),
        "roofline": roofline,
    }
    if args.prefetch and w.name == "c2":
        out["prefetch"] = prefetch_leg(ctx, w, cams, args, d_rgba, d_hits, frame_ms, uploads, U, passes)
    if not args.no_cpu:
        log("cpu baseline")
        cb, objs = cpu_baseline(w, cams[args.warmup: args.warmup + args.steps], args.cpu_seconds)
        if w.name == "c1":   # the per-core figure: one thread on the same frames
            one, _ = cpu_baseline(w, cams[args.warmup: args.warmup + args.steps], min(args.cpu_seconds, 6.0),
                                  threads=1)
            cb["one_thread"] = {"value": one["value"], "sample": one["sample"]}
        out["cpu_baseline"] = cb
        log("parity sample")
        out["parity"] = parity_sample(w, ctx, cams[-1], objs, args.parity_rays, seed=1)
    else:
        out["cpu_baseline"] = None
    d_rgba.free()
    d_hits.free()
    ctx.close()
    if getattr(w, "tmpdir", None):
        shutil.rmtree(w.tmpdir, ignore_errors=True)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
