"""Shared scene setup for the GPU parity tests (oracle = the checker)."""
import math

import numpy as np

from conftest import blobs_volume


def flat_runs(runs, max_level):
    return np.concatenate(runs[: max_level + 1]).astype(np.uint64), [r.size for r in runs[: max_level + 1]]


class Scene:
    def __init__(self, vto, vol, iso, level_o, level_c, literal=False):
        self.vol = np.ascontiguousarray(vol, np.float32)
        self.iso = iso
        self.level_o, self.level_c = level_o, level_c
        self.n, self.runs = vto.build_octree(self.vol, iso, literal=literal)
        self.cube = 1 << (self.n - level_c)
        self.V = vto.Volume(self.vol)
        self.O = vto.Octree(self.runs, self.n, level_o)
        self.prm_o = vto.params(level_o, level_c, iso)

    def setup_gpu(self, vt, ctx, slots, inc=3):
        flat, sizes = flat_runs(self.runs, self.level_o)
        ctx.octree_upload(flat, sizes, self.n, self.level_o)
        self.d_vol = ctx.to_device(self.vol)
        ctx.cache_create(self.level_c, self.cube, inc, slots, 0)
        ctx.cache_attach_device_volume(self.d_vol, self.vol.shape)
        self.prm_g = vt.render_params(self.level_o, self.level_c, self.iso)

    def render_gpu(self, vt, ctx, cam):
        R = cam.W * cam.H
        rgba = ctx.alloc(R * 16)
        hits = ctx.alloc(R * 32)
        try:
            st = ctx.render_frame(cam, self.prm_g, rgba, hits)
            return rgba.download(np.float32, (R, 4)), hits.download(vt.HIT_DTYPE, (R,)), st
        finally:
            rgba.free()
            hits.free()

    def render_cpu(self, vto, cam, nthreads=0):
        rgba, hits, _ = vto.render(self.V, self.O, cam, self.prm_o, nthreads=nthreads)
        return rgba, hits


def orbit_camera(vt, W, H, center, radius, theta_deg, phi_deg=20.0, fov=45.0):
    t, p = math.radians(theta_deg), math.radians(phi_deg)
    eye = np.array(center, np.float64) + radius * np.array([math.sin(t) * math.cos(p), math.sin(p),
                                                            math.cos(t) * math.cos(p)])
    look, up, right = vt.look_at(eye, center)
    return vt.make_camera(W, H, eye, look, up, right, fov_v_deg=fov)


def assert_same(hits_g, rgba_g, hits_c, rgba_c, depth_tol=1e-4, rgba_tol=1.0 / 255.0):
    """The parity bar: node, brick and hit voxel bit-exact; depth within
    1e-4 voxel; RGBA within 1/255 (NaN matches NaN)."""
    bad = np.nonzero(hits_g["node"] != hits_c["node"])[0]
    assert bad.size == 0, f"node mismatch at rays {bad[:10]}: {hits_g['node'][bad[:5]]} vs {hits_c['node'][bad[:5]]}"
    assert np.array_equal(hits_g["brick"], hits_c["brick"])
    assert np.array_equal(hits_g["voxel"], hits_c["voxel"])
    dg, dc = hits_g["depth"].astype(np.float64), hits_c["depth"].astype(np.float64)
    nan = np.isnan(dc)
    assert np.array_equal(np.isnan(dg), nan)
    assert np.all(np.abs(dg[~nan] - dc[~nan]) <= depth_tol)
    ng = np.isnan(rgba_g)
    assert np.array_equal(ng, np.isnan(rgba_c))
    assert np.all(np.abs(rgba_g[~ng].astype(np.float64) - rgba_c[~ng]) <= rgba_tol)


def assert_bits(hits_g, rgba_g, hits_c, rgba_c):
    """Bit-identical hit records and pixels, NaN matching NaN (a 0/0 secant
    gives NaN on both sides, with a sign bit that differs between the GPU and
    the host)."""
    for f in ("node", "brick", "voxel"):
        assert np.array_equal(hits_g[f], hits_c[f]), f
    dg, dc = hits_g["depth"], hits_c["depth"]
    nd = np.isnan(dg)
    assert np.array_equal(nd, np.isnan(dc))
    assert np.array_equal(dg[~nd].view(np.uint32), dc[~nd].view(np.uint32))
    nr = np.isnan(rgba_g)
    assert np.array_equal(nr, np.isnan(rgba_c))
    assert np.array_equal(rgba_g[~nr].view(np.uint32), rgba_c[~nr].view(np.uint32))


def bitexact_fraction(hits_g, rgba_g, hits_c, rgba_c):
    same = (hits_g["depth"].view(np.uint32) == hits_c["depth"].view(np.uint32)) & \
           np.all(rgba_g.view(np.uint32) == rgba_c.view(np.uint32), axis=1)
    return float(same.mean())


__all__ = ["Scene", "orbit_camera", "assert_same", "assert_bits", "bitexact_fraction", "blobs_volume", "flat_runs"]
