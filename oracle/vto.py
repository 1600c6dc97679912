"""ctypes wrapper of oracle/_build/libvt_oracle.so -- test infrastructure only.

Each function forwards to the C restatement in vt_oracle.c, which cites the
reference lines it follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libvt_oracle.so"
_P = ctypes.c_void_p
_lib = None


class OCamera(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("look", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("right", ctypes.c_float * 3), ("h", ctypes.c_float), ("w", ctypes.c_float),
                ("distance", ctypes.c_float), ("W", ctypes.c_int32), ("H", ctypes.c_int32)]


class OVolume(ctypes.Structure):
    _fields_ = [("data", _P), ("dims", ctypes.c_int32 * 3), ("origin", ctypes.c_int32 * 3)]


class OOctree(ctypes.Structure):
    _fields_ = [("nLevels", ctypes.c_int32), ("maxLevel", ctypes.c_int32), ("runs", _P), ("sizes", _P)]


class OParams(ctypes.Structure):
    _fields_ = [("levelO", ctypes.c_int32), ("levelC", ctypes.c_int32), ("iso", ctypes.c_float),
                ("light", ctypes.c_float * 3), ("color_scale", ctypes.c_float)]


HIT_DTYPE = np.dtype([("node", "<u8"), ("brick", "<u8"), ("voxel", "<i4", (3,)), ("depth", "<f4")])

_U64, _I32, _F = ctypes.c_uint64, ctypes.c_int32, ctypes.c_float
_SIGS = {
    "vto_dilate": (_U64, [_U64]),
    "vto_encode": (_U64, [ctypes.c_int] * 5),
    "vto_level": (ctypes.c_int, [_U64]),
    "vto_decode": (None, [_U64, ctypes.c_int, ctypes.c_int, _P]),
    "vto_update_coordinates": (None, [ctypes.c_int, ctypes.c_int, _U64, ctypes.c_int, _U64, _P, _P]),
    "vto_children_mask_device": (ctypes.c_uint32, [_P, ctypes.c_int, _U64]),
    "vto_children_mask_builder": (ctypes.c_uint32, [_P, ctypes.c_int, _U64]),
    "vto_member_exact": (ctypes.c_int, [_P, ctypes.c_int, _U64]),
    "vto_slab": (ctypes.c_int, [_P, _P, _P, _P, ctypes.POINTER(_F), ctypes.POINTER(_F)]),
    "vto_screen_size": (None, [_F, _F, _F, ctypes.POINTER(_F), ctypes.POINTER(_F)]),
    "vto_camera_move": (None, [ctypes.POINTER(OCamera), ctypes.c_int, _P]),
    "vto_ray_dir": (None, [ctypes.POINTER(OCamera), ctypes.c_int, ctypes.c_int, _P]),
    "vto_voxel": (_F, [ctypes.POINTER(OVolume), ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "vto_sample": (_F, [ctypes.POINTER(OVolume), _F, _F, _F]),
    "vto_read_box": (None, [ctypes.POINTER(OVolume), _P, _P, _P]),
    "vto_traverse": (ctypes.c_int, [ctypes.POINTER(OOctree), _P, _P, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int]),
    "vto_render": (ctypes.c_int, [ctypes.POINTER(OVolume), ctypes.POINTER(OOctree), ctypes.POINTER(OCamera),
                                  ctypes.POINTER(OParams), ctypes.c_int64, ctypes.c_int64, _P, _P, _P, ctypes.c_int]),
    "vto_render_list": (ctypes.c_int, [ctypes.POINTER(OVolume), ctypes.POINTER(OOctree), ctypes.POINTER(OCamera),
                                       ctypes.POINTER(OParams), _P, ctypes.c_int64, _P, _P, ctypes.c_int]),
    "vto_build_octree": (ctypes.c_int, [ctypes.POINTER(OVolume), _F, ctypes.c_int, ctypes.POINTER(_P), _P,
                                        ctypes.c_int]),
    "vto_free": (None, [_P]),
    "vto_render_cube": (ctypes.c_int, [ctypes.POINTER(OVolume), ctypes.POINTER(OCamera), _F, _P, _P, ctypes.c_int, _P]),
    "vto_render_cached": (ctypes.c_int, [ctypes.POINTER(OOctree), ctypes.POINTER(OCamera), ctypes.POINTER(OParams),
                                         _P, _P, _P, _P, _P, _P, ctypes.c_int]),
    "vto_cache_new": (_P, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64]),
    "vto_cache_free": (None, [_P]),
    "vto_h5_open": (_P, [ctypes.c_char_p, ctypes.c_char_p]),
    "vto_h5_close": (None, [_P]),
    "vto_octree_write": (ctypes.c_int, [ctypes.c_char_p, _F, _I32, _P, _I32, _P, _P]),
    "vto_octree_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_F), ctypes.POINTER(_I32), _P,
                                       ctypes.POINTER(_I32), ctypes.POINTER(_P), _P, ctypes.c_int]),
}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _f3(v):
    return np.ascontiguousarray(v, np.float32)


def _i3(v):
    return np.ascontiguousarray(v, np.int32)


# ---------------------------------------------------------------- Morton
def encode(x, y, z, level, n):
    return int(lib().vto_encode(int(x), int(y), int(z), int(level), int(n)))


def decode(code, level, n):
    out = np.zeros(3, np.int32)
    lib().vto_decode(int(code), int(level), int(n), out.ctypes.data)
    return tuple(int(v) for v in out)


def level(code):
    return int(lib().vto_level(int(code)))


def dilate(x):
    return int(lib().vto_dilate(int(x)))


def update_coordinates(max_level, c_level, c_index, n_level, n_index, min_box):
    mb = _i3(min_box)
    out = np.zeros(3, np.int32)
    lib().vto_update_coordinates(max_level, c_level, int(c_index), n_level, int(n_index), mb.ctypes.data,
                                 out.ctypes.data)
    return tuple(int(v) for v in out)


# -------------------------------------------------------------- run lists
def children_mask_device(runs, parent):
    r = np.ascontiguousarray(runs, np.uint64)
    return int(lib().vto_children_mask_device(r.ctypes.data, r.size, int(parent)))


def children_mask_builder(runs, parent):
    r = np.ascontiguousarray(runs, np.uint64)
    return int(lib().vto_children_mask_builder(r.ctypes.data, r.size, int(parent)))


def member_exact(runs, code):
    r = np.ascontiguousarray(runs, np.uint64)
    return bool(lib().vto_member_exact(r.ctypes.data, r.size, int(code)))


# ------------------------------------------------------------------- slab
def slab(o, d, mn, mx):
    o, d, mn, mx = _f3(o), _f3(d), _i3(mn), _i3(mx)
    tn, tf = ctypes.c_float(), ctypes.c_float()
    hit = lib().vto_slab(o.ctypes.data, d.ctypes.data, mn.ctypes.data, mx.ctypes.data, ctypes.byref(tn),
                         ctypes.byref(tf))
    return bool(hit), tn.value, tf.value


# ----------------------------------------------------------------- camera
def camera_from(cam) -> OCamera:
    """Copy a visualtur_amd.Camera (same field layout) or dict into an OCamera."""
    oc = OCamera()
    if isinstance(cam, dict):
        for k, v in cam.items():
            if k in ("pos", "look", "up", "right"):
                getattr(oc, k)[:] = [float(x) for x in v]
            else:
                setattr(oc, k, v)
    else:
        ctypes.memmove(ctypes.byref(oc), ctypes.byref(cam), ctypes.sizeof(OCamera))
    return oc


CAMERA_MOVES = {"Move": 0, "RotateX": 1, "RotateY": 2, "RotateZ": 3, "MoveForward": 4, "MoveUpward": 5,
                "StrafeRight": 6}


def camera_move(cam: OCamera, op: str, arg) -> OCamera:
    """Camera.cu:83-156 applied to `cam` in place (and returned); `arg` is a
    3-vector for Move, an angle in degrees or a distance otherwise."""
    a = np.ascontiguousarray(np.atleast_1d(np.asarray(arg, np.float32)), np.float32)
    lib().vto_camera_move(ctypes.byref(cam), CAMERA_MOVES[op], a.ctypes.data)
    return cam


def camera_vectors(cam: OCamera) -> dict:
    return {k: np.array(getattr(cam, k)[:], np.float32) for k in ("pos", "look", "up", "right")}


def screen_size(distance, fov_h, fov_w):
    h, w = ctypes.c_float(), ctypes.c_float()
    lib().vto_screen_size(distance, fov_h, fov_w, ctypes.byref(h), ctypes.byref(w))
    return h.value, w.value


def ray_dirs(cam) -> np.ndarray:
    oc = camera_from(cam)
    out = np.empty((oc.W * oc.H, 3), np.float32)
    tmp = np.empty(3, np.float32)
    for i in range(oc.H):
        for j in range(oc.W):
            lib().vto_ray_dir(ctypes.byref(oc), i, j, tmp.ctypes.data)
            out[i * oc.W + j] = tmp
    return out


# ----------------------------------------------------------------- volume
class Volume:
    def __init__(self, arr: np.ndarray):
        self.arr = np.ascontiguousarray(arr, np.float32)
        self.c = OVolume()
        self.c.data = self.arr.ctypes.data
        self.c.dims[:] = list(self.arr.shape)

    def sample(self, p):
        return float(lib().vto_sample(ctypes.byref(self.c), float(p[0]), float(p[1]), float(p[2])))

    def voxel(self, x, y, z):
        return float(lib().vto_voxel(ctypes.byref(self.c), int(x), int(y), int(z)))

    def read_box(self, s, e):
        s3, e3 = _i3(s), _i3(e)
        out = np.empty(tuple(abs(int(e3[k]) - int(s3[k])) for k in range(3)), np.float32)
        lib().vto_read_box(ctypes.byref(self.c), s3.ctypes.data, e3.ctypes.data, out.ctypes.data)
        return out


# ----------------------------------------------------------------- octree
class Octree:
    """Run lists of levels 0..maxLevel (flattened as in Octree.cu:820-839)."""

    def __init__(self, runs_by_level, n_levels: int, max_level: int | None = None):
        if max_level is None:
            max_level = len(runs_by_level) - 1
        self.levels = [np.ascontiguousarray(r, np.uint64) for r in runs_by_level]
        self.n = n_levels
        self.max_level = max_level
        self.flat = np.ascontiguousarray(np.concatenate(self.levels[: max_level + 1]) if max_level >= 0
                                         else np.zeros(0, np.uint64), np.uint64)
        self.sizes = np.array([r.size for r in self.levels[: max_level + 1]], np.int32)
        self.c = OOctree()
        self.c.nLevels = n_levels
        self.c.maxLevel = max_level
        self.c.runs = self.flat.ctypes.data
        self.c.sizes = self.sizes.ctypes.data

    def traverse(self, o, d, final_level, max_out=64, cap_total=1 << 20):
        out = np.zeros(max(max_out, 1), np.uint64)
        o, d = _f3(o), _f3(d)
        cnt = lib().vto_traverse(ctypes.byref(self.c), o.ctypes.data, d.ctypes.data, final_level, out.ctypes.data,
                                 max_out, cap_total)
        return [int(v) for v in out[: min(cnt, max_out)]], int(cnt)


def build_octree(vol: np.ndarray, iso: float, literal: bool = True):
    """CreateOctree restatement: returns (n_levels, [runs level 0..n])."""
    v = Volume(vol)
    sizes = np.zeros(32, np.int32)
    runs_p = ctypes.c_void_p()
    n = lib().vto_build_octree(ctypes.byref(v.c), float(iso), 1 if literal else 0, ctypes.byref(runs_p),
                               sizes.ctypes.data, 32)
    if n < 0:
        raise ValueError("oracle octree build failed")
    total = int(sizes[: n + 1].sum())
    flat = np.ctypeslib.as_array((ctypes.c_uint64 * max(total, 1)).from_address(runs_p.value)).copy()[:total]
    lib().vto_free(runs_p)
    out, off = [], 0
    for l in range(n + 1):
        out.append(flat[off: off + sizes[l]].copy())
        off += sizes[l]
    return n, out


def octree_write(path, iso, dimension, real_dim, n, runs_by_level):
    flat = np.ascontiguousarray(np.concatenate(runs_by_level), np.uint64)
    sizes = np.array([r.size for r in runs_by_level], np.int32)
    rd = _i3(real_dim)
    rc = lib().vto_octree_write(str(path).encode(), float(iso), int(dimension), rd.ctypes.data, int(n),
                                flat.ctypes.data, sizes.ctypes.data)
    if rc != 0:
        raise OSError(f"octree write failed: {path}")


def octree_read(path):
    iso, dim, n = ctypes.c_float(), ctypes.c_int32(), ctypes.c_int32()
    rd = np.zeros(3, np.int32)
    sizes = np.zeros(32, np.int32)
    runs_p = ctypes.c_void_p()
    rc = lib().vto_octree_read(str(path).encode(), ctypes.byref(iso), ctypes.byref(dim), rd.ctypes.data,
                               ctypes.byref(n), ctypes.byref(runs_p), sizes.ctypes.data, 32)
    if rc != 0:
        raise ValueError(f"octree read failed ({rc}): {path}")
    nl = n.value
    total = int(sizes[: nl + 1].sum())
    flat = np.ctypeslib.as_array((ctypes.c_uint64 * max(total, 1)).from_address(runs_p.value)).copy()[:total]
    lib().vto_free(runs_p)
    out, off = [], 0
    for l in range(nl + 1):
        out.append(flat[off: off + sizes[l]].copy())
        off += sizes[l]
    return iso.value, dim.value, tuple(int(v) for v in rd), nl, out


# ------------------------------------------------------------------ frame
def params(level_o, level_c, iso, light=(0.0, 512.0, 0.0), color_scale=256.0) -> OParams:
    p = OParams()
    p.levelO, p.levelC, p.iso = level_o, level_c, iso
    p.light[:] = [float(v) for v in light]
    p.color_scale = color_scale
    return p


def render(vol: Volume, oct: Octree, cam, prm: OParams, ray0=0, ray1=None, nthreads=0, trace=False):
    """Full-frame (or [ray0, ray1) band) render; returns rgba[R,4], hits[R], brick_seen."""
    oc = camera_from(cam)
    R = oc.W * oc.H
    if ray1 is None:
        ray1 = R
    rgba = np.zeros((R, 4), np.float32)
    hits = np.zeros(R, HIT_DTYPE)
    seen = np.zeros(1 << (3 * prm.levelC), np.uint8) if trace else None
    rc = lib().vto_render(ctypes.byref(vol.c), ctypes.byref(oct.c), ctypes.byref(oc), ctypes.byref(prm), ray0, ray1,
                          rgba.ctypes.data, hits.ctypes.data, seen.ctypes.data if trace else None, nthreads)
    if rc != 0:
        raise ValueError("oracle render: bad parameters")
    return rgba, hits, seen


def render_list(vol: Volume, oct: Octree, cam, prm: OParams, ids, nthreads=0):
    oc = camera_from(cam)
    ids = np.ascontiguousarray(ids, np.int64)
    rgba = np.zeros((ids.size, 4), np.float32)
    hits = np.zeros(ids.size, HIT_DTYPE)
    rc = lib().vto_render_list(ctypes.byref(vol.c), ctypes.byref(oct.c), ctypes.byref(oc), ctypes.byref(prm),
                               ids.ctypes.data, ids.size, rgba.ctypes.data, hits.ctypes.data, nthreads)
    if rc != 0:
        raise ValueError("oracle render: bad parameters")
    return rgba, hits


class CacheStats(ctypes.Structure):
    _fields_ = [("passes", ctypes.c_int64), ("hits", ctypes.c_int64), ("misses", ctypes.c_int64),
                ("evictions", ctypes.c_int64)]


class H5Source:
    """An HDF5 dataset as the cached render's brick source (FileManager's box read)."""

    def __init__(self, path, dataset="value"):
        self.h = lib().vto_h5_open(str(path).encode(), dataset.encode())
        if not self.h:
            raise OSError(f"oracle: cannot open {path}:{dataset}")

    def close(self):
        if self.h:
            lib().vto_h5_close(self.h)
            self.h = None


class BrickLRU:
    """The cached render's LRU brick cache, kept across frames."""

    def __init__(self, level_c: int, cube: int, inc: int, capacity: int):
        self.h = lib().vto_cache_new(level_c, cube, inc, capacity)
        if not self.h:
            raise MemoryError("oracle: brick cache")

    def close(self):
        if self.h:
            lib().vto_cache_free(self.h)
            self.h = None


def render_cached(oct: Octree, cam, prm: OParams, cache: BrickLRU, source, nthreads=0):
    """The reference frame loop through `cache` fed from `source` (a Volume
    or an H5Source); returns rgba[R,4], hits[R], stats."""
    oc = camera_from(cam)
    R = oc.W * oc.H
    rgba = np.zeros((R, 4), np.float32)
    hits = np.zeros(R, HIT_DTYPE)
    st = CacheStats()
    if isinstance(source, H5Source):
        fn, user = ctypes.cast(lib().vto_h5_box_reader, _P), source.h
    else:
        fn, user = ctypes.cast(lib().vto_volume_box_reader, _P), ctypes.addressof(source.c)
    rc = lib().vto_render_cached(ctypes.byref(oct.c), ctypes.byref(oc), ctypes.byref(prm), cache.h, fn, user,
                                 rgba.ctypes.data, hits.ctypes.data, ctypes.byref(st), nthreads)
    if rc != 0:
        raise ValueError(f"oracle render_cached: status {rc}")
    return rgba, hits, {k: getattr(st, k) for k, _ in st._fields_}


def render_cube(vol: Volume, cam, iso, light, min_box, dim):
    """rayCaster::renderCube restatement; returns rgba[R,4]."""
    oc = camera_from(cam)
    rgba = np.zeros((oc.W * oc.H, 4), np.float32)
    li, mb = _f3(light), _i3(min_box)
    lib().vto_render_cube(ctypes.byref(vol.c), ctypes.byref(oc), float(iso), li.ctypes.data, mb.ctypes.data,
                          int(dim), rgba.ctypes.data)
    return rgba
