"""Learned brick-visibility prefetching: camera paths, visited-brick traces,
training of the MLP on the GPU, and the plain-vs-prefetch comparison.

The plain renderer records which bricks each frame marches (the trace).  The
MLP (vt_mlp_*, BF16 MFMA kernels) learns P(brick marched | camera, brick)
from the traces of training paths.  Before a frame, vt_prefetch scores every
candidate brick for the upcoming camera and loads the likely ones.  The
renderer still decides every brick it marches, so the pictures and hit
records are the same with or without the prefetcher.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import Context, MLP, look_at, make_camera


def random_walk(W, H, D, frames, seed, radius=(1.6, 2.6), fov=45.0):
    """Seeded random walk of a camera looking near the volume centre."""
    rng = np.random.default_rng(seed)
    centre = np.array([D / 2.0] * 3)
    theta = rng.uniform(0.0, 360.0)
    phi = rng.uniform(-30.0, 30.0)
    r = rng.uniform(*radius)
    jitter = np.zeros(3)
    cams = []
    for _ in range(frames):
        theta += rng.normal(0.0, 4.0)
        phi = float(np.clip(phi + rng.normal(0.0, 2.0), -60.0, 60.0))
        r = float(np.clip(r + rng.normal(0.0, 0.03), *radius))
        jitter = np.clip(jitter + rng.normal(0.0, 0.01 * D, 3), -0.1 * D, 0.1 * D)
        t, p = math.radians(theta), math.radians(phi)
        eye = centre + r * D * np.array([math.sin(t) * math.cos(p), math.sin(p), math.cos(t) * math.cos(p)])
        look, up, right = look_at(eye, centre + jitter)
        cams.append(make_camera(W, H, eye, look, up, right, fov_v_deg=fov))
    return cams


def candidate_codes(runs_level_c) -> np.ndarray:
    """Codes of the bricks that hold octree nodes (the level-levelC run list)."""
    r = np.asarray(runs_level_c, np.uint64)
    parts = [np.arange(int(r[j]), int(r[j + 1]) + 1, dtype=np.uint64) for j in range(0, r.size, 2)]
    return np.concatenate(parts) if parts else np.zeros(0, np.uint64)


def collect_traces(ctx: Context, cams, prm, level_c: int):
    """Render each camera with the visited-brick trace on; returns, per frame,
    the level-levelC codes of the bricks marched."""
    R = cams[0].W * cams[0].H
    d_rgba, d_hits = ctx.alloc(R * 16), ctx.alloc(R * 32)
    base = 1 << (3 * level_c)
    out = []
    ctx.set_frame_options(collect_stats=False, record_trace=True)
    try:
        for cam in cams:
            ctx.render_frame(cam, prm, d_rgba, d_hits)
            out.append(np.nonzero(ctx.brick_trace(level_c))[0].astype(np.uint64) + np.uint64(base))
    finally:
        ctx.set_frame_options(False, False)
        d_rgba.free()
        d_hits.free()
    return out


def build_dataset(ctx: Context, net: MLP, cams, traces, candidates, n_levels, level_c, neg_per_pos=4, seed=0,
                  max_pos=None):
    """Features (computed by vt_mlp_features) and labels: every visited brick
    (or max_pos of them, sampled, per frame) is a positive, plus neg_per_pos
    sampled non-visited candidates each."""
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    d_codes = ctx.alloc(max(1, candidates.size) * 8)
    d_x = ctx.alloc(max(1, candidates.size) * 48)
    try:
        for cam, pos in zip(cams, traces):
            neg_pool = np.setdiff1d(candidates, pos, assume_unique=False)
            if max_pos is not None and pos.size > max_pos:
                pos = rng.choice(pos, size=max_pos, replace=False)
            k = min(neg_pool.size, max(8, neg_per_pos * pos.size))
            neg = rng.choice(neg_pool, size=k, replace=False) if k else np.zeros(0, np.uint64)
            codes = np.concatenate([pos, neg]).astype(np.uint64)
            labels = np.concatenate([np.ones(pos.size, np.float32), np.zeros(neg.size, np.float32)])
            for o in range(0, codes.size, net.max_batch):
                c = codes[o: o + net.max_batch]
                d_codes.upload(c)
                net.features(cam, n_levels, level_c, d_codes, c.size, d_x)
                xs.append(d_x.download(np.float32, (c.size, 12)))
            ys.append(labels)
    finally:
        d_codes.free()
        d_x.free()
    return np.concatenate(xs), np.concatenate(ys)


def train(ctx: Context, net: MLP, X, Y, steps=400, batch=4096, lr=1e-3, seed=0):
    """Adam on shuffled mini-batches, entirely in the MFMA kernels; returns the
    loss curve.  Positives are weighted by the negative/positive ratio."""
    rng = np.random.default_rng(seed)
    pos_weight = float((Y < 0.5).sum()) / max(1.0, float((Y > 0.5).sum()))
    batch = min(batch, net.max_batch, X.shape[0])
    d_x, d_y = ctx.alloc(batch * 48), ctx.alloc(batch * 4)
    losses = []
    try:
        order = rng.permutation(X.shape[0])
        pos = 0
        for _ in range(steps):
            if pos + batch > order.size:
                order = rng.permutation(X.shape[0])
                pos = 0
            idx = order[pos: pos + batch]
            pos += batch
            d_x.upload(np.ascontiguousarray(X[idx]))
            d_y.upload(np.ascontiguousarray(Y[idx]))
            losses.append(net.train_step(d_x, d_y, batch, lr, pos_weight))
    finally:
        d_x.free()
        d_y.free()
    return np.array(losses), pos_weight


def run_path(ctx: Context, cams, prm, net: MLP | None = None, budget=0, min_prob=0.5):
    """Render a camera path from the current cache state; with `net`, prefetch
    before every frame.  Returns per-frame statistics."""
    R = cams[0].W * cams[0].H
    d_rgba, d_hits = ctx.alloc(R * 16), ctx.alloc(R * 32)
    rows = []
    try:
        for cam in cams:
            t0 = time.perf_counter()
            pre = net.prefetch(cam, budget, min_prob) if net is not None else 0
            t1 = time.perf_counter()
            st = ctx.render_frame(cam, prm, d_rgba, d_hits)
            t2 = time.perf_counter()
            rows.append({"prefetched": pre, "demand_misses": st.bricks_uploaded, "passes": st.passes,
                         "ms_prefetch": (t1 - t0) * 1e3, "ms_frame": (t2 - t1) * 1e3,
                         "ms_total": (t2 - t0) * 1e3})
        last = (d_rgba.download(np.float32, (R, 4)), d_hits.download(np.dtype([("node", "<u8"), ("brick", "<u8"),
                                                                               ("voxel", "<i4", (3,)),
                                                                               ("depth", "<f4")]), (R,)))
    finally:
        d_rgba.free()
        d_hits.free()
    return rows, last
