"""visualtur_amd -- MI355X isosurface ray casting (the visualTur hot path).

Python access to the C ABI of include/visualtur_amd.h through ctypes.  The
library is the HIP build in visualtur_amd/_build/libvt_hip.so; there is no
fallback: if it is missing or no GPU is present, the calls raise.
"""
from __future__ import annotations

import ctypes
import math
from pathlib import Path

import numpy as np

__all__ = [
    "VTError", "Context", "DeviceBuffer", "Camera", "RenderParams", "FrameStats", "CacheStats",
    "HIT_DTYPE", "lib", "make_camera", "look_at", "screen_size", "file_create", "File",
]

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "_build" / "libvt_hip.so"
_lib = None


class VTError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Camera(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("look", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("right", ctypes.c_float * 3), ("h", ctypes.c_float), ("w", ctypes.c_float),
                ("distance", ctypes.c_float), ("W", ctypes.c_int32), ("H", ctypes.c_int32)]


class RenderParams(ctypes.Structure):
    _fields_ = [("levelO", ctypes.c_int32), ("levelC", ctypes.c_int32), ("iso", ctypes.c_float),
                ("light", ctypes.c_float * 3), ("color_scale", ctypes.c_float)]


class FrameStats(ctypes.Structure):
    _fields_ = [("passes", ctypes.c_int32), ("err", ctypes.c_int32), ("bricks_requested", ctypes.c_int64),
                ("bricks_uploaded", ctypes.c_int64), ("host_tier_misses", ctypes.c_int64),
                ("rays_parked", ctypes.c_int64), ("samples", ctypes.c_uint64), ("nodes", ctypes.c_uint64),
                ("ms_total", ctypes.c_double), ("ms_kernel", ctypes.c_double), ("ms_service", ctypes.c_double),
                ("bricks_early", ctypes.c_int64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class CacheStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("gpu_slots", "gpu_resident", "cpu_slots", "cpu_resident",
                                             "lookups_missed", "uploads", "file_reads", "evictions")]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class GenParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("seed", ctypes.c_uint32), ("n", ctypes.c_int32),
                ("count", ctypes.c_int32), ("shapes", ctypes.c_void_p), ("octaves", ctypes.c_int32),
                ("base_freq", ctypes.c_float), ("gain", ctypes.c_float), ("lacunarity", ctypes.c_float),
                ("scale", ctypes.c_float)]


GEN_BLOBS, GEN_FBM, GEN_SDF, GEN_FOURIER = 0, 1, 2, 3

class Request(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_int32), ("cam", Camera), ("iso", ctypes.c_float), ("light", ctypes.c_float * 3),
                ("color_scale", ctypes.c_float), ("arrival_s", ctypes.c_double), ("h_rgba", ctypes.c_void_p),
                ("h_hits", ctypes.c_void_p)]


class RequestResult(ctypes.Structure):
    _fields_ = [("admit_s", ctypes.c_double), ("done_s", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("launches", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class ServeStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_int64), ("launches", ctypes.c_int64), ("tiles", ctypes.c_int64),
                ("rays", ctypes.c_int64), ("wall_s", ctypes.c_double), ("busy_s", ctypes.c_double),
                ("tables_built", ctypes.c_int64), ("tables_held", ctypes.c_int64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class FramesStats(ctypes.Structure):
    _fields_ = [("frames", ctypes.c_int64), ("launches", ctypes.c_int64), ("bricks_requested", ctypes.c_int64),
                ("bricks_uploaded", ctypes.c_int64), ("bricks_early", ctypes.c_int64),
                ("bricks_prefetched", ctypes.c_int64), ("host_tier_misses", ctypes.c_int64),
                ("ms_kernel", ctypes.c_double), ("wall_s", ctypes.c_double), ("busy_s", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


HIT_DTYPE = np.dtype([("node", "<u8"), ("brick", "<u8"), ("voxel", "<i4", (3,)), ("depth", "<f4")])
assert HIT_DTYPE.itemsize == 32

_P = ctypes.c_void_p
_SIGS = {
    "vt_last_error": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_size_t]),
    "vt_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    "vt_context_destroy": (ctypes.c_int, [_P]),
    "vt_synchronize": (ctypes.c_int, [_P]),
    "vt_malloc": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    "vt_free": (ctypes.c_int, [_P, _P]),
    "vt_memcpy_h2d": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "vt_memcpy_d2h": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "vt_memset": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_size_t]),
    "vt_octree_upload": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32, ctypes.c_int32]),
    "vt_octree_load_file": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "vt_camera_rays": (ctypes.c_int, [_P, ctypes.POINTER(Camera), _P]),
    "vt_screen_size": (ctypes.c_int, [ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                      ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "vt_cache_create": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                       ctypes.c_int64]),
    "vt_cache_attach_hdf5": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_char_p]),
    "vt_cache_attach_device_volume": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "vt_cache_prefetch_host": (ctypes.c_int, [_P, ctypes.c_int]),
    "vt_cache_load_bricks": (ctypes.c_int, [_P, _P, ctypes.c_int64]),
    "vt_cache_read_brick": (ctypes.c_int, [_P, ctypes.c_uint64, _P]),
    "vt_cache_clear": (ctypes.c_int, [_P]),
    "vt_cache_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(CacheStats)]),
    "vt_cache_destroy": (ctypes.c_int, [_P]),
    "vt_render_frame": (ctypes.c_int, [_P, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams), _P, _P,
                                       ctypes.POINTER(FrameStats)]),
    "vt_render_frame_noalloc": (ctypes.c_int, [_P, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams), _P, _P,
                                               ctypes.POINTER(FrameStats)]),
    "vt_set_frame_options": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    "vt_get_brick_trace": (ctypes.c_int, [_P, _P, ctypes.c_int64]),
    "vt_set_march_options": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "vt_set_wide_march": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int64]),
    "vt_set_early_loads": (ctypes.c_int, [_P, ctypes.c_int64]),
    "vt_set_first_pass_budget": (ctypes.c_int, [_P, ctypes.c_int32]),
    "vt_set_first_pass_tiles": (ctypes.c_int, [_P, ctypes.c_int32]),
    "vt_trace_nodes": (ctypes.c_int, [_P, ctypes.POINTER(Camera), ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int32, _P, _P]),
    "vt_octree_build": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _P]),
    "vt_octree_get_runs": (ctypes.c_int, [_P, ctypes.c_int32, _P, ctypes.c_int64]),
    "vt_octree_write_file": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "vt_volume_generate": (ctypes.c_int, [_P, ctypes.POINTER(GenParams), ctypes.c_int64, ctypes.c_int64,
                                          ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _P]),
    "vt_volume_write_hdf5": (ctypes.c_int, [_P, _P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_char_p,
                                            ctypes.c_char_p, ctypes.c_int32]),
    "vt_image_write_ppm": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p]),
    "vt_copy_probe": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "vt_frames_create": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(_P)]),
    "vt_frames_set_prefetch": (ctypes.c_int, [_P, _P, ctypes.c_int64, ctypes.c_float]),
    "vt_frames_submit": (ctypes.c_int, [_P, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams), _P, _P,
                                        ctypes.POINTER(ctypes.c_int64)]),
    "vt_frames_poll": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(FrameStats)]),
    "vt_frames_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(FramesStats)]),
    "vt_frames_destroy": (ctypes.c_int, [_P]),
    "vt_cache_attach_host_from_device": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "vt_server_create": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_P)]),
    "vt_server_destroy": (ctypes.c_int, [_P]),
    "vt_server_set_lanes": (ctypes.c_int, [_P, ctypes.c_int64]),
    "vt_server_add_scene": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "vt_server_add_scene_range": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                                 ctypes.c_float, ctypes.POINTER(ctypes.c_int32),
                                                 ctypes.POINTER(ctypes.c_int64)]),
    "vt_server_run": (ctypes.c_int, [_P, _P, ctypes.c_int64, ctypes.c_int32, _P, ctypes.POINTER(ServeStats)]),
    "vt_server_start": (ctypes.c_int, [_P, ctypes.c_int32]),
    "vt_server_stop": (ctypes.c_int, [_P, ctypes.POINTER(ServeStats)]),
    "vt_server_submit": (ctypes.c_int, [_P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_int64)]),
    "vt_server_poll": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(RequestResult)]),
    "vt_mlp_num_params": (ctypes.c_int, []),
    "vt_prefetch": (ctypes.c_int, [_P, _P, ctypes.POINTER(Camera), ctypes.c_int64, ctypes.c_float,
                                   ctypes.POINTER(ctypes.c_int64)]),
    "vt_mlp_create": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(_P)]),
    "vt_mlp_destroy": (ctypes.c_int, [_P]),
    "vt_mlp_get_params": (ctypes.c_int, [_P, _P]),
    "vt_mlp_set_params": (ctypes.c_int, [_P, _P]),
    "vt_mlp_forward": (ctypes.c_int, [_P, _P, ctypes.c_int64, _P]),
    "vt_mlp_gradients": (ctypes.c_int, [_P, _P, _P, ctypes.c_int64, ctypes.c_float, _P,
                                        ctypes.POINTER(ctypes.c_float)]),
    "vt_mlp_train_step": (ctypes.c_int, [_P, _P, _P, ctypes.c_int64, ctypes.c_float, ctypes.c_float,
                                         ctypes.POINTER(ctypes.c_float)]),
    "vt_mlp_features": (ctypes.c_int, [_P, ctypes.POINTER(Camera), ctypes.c_int32, ctypes.c_int32, _P,
                                       ctypes.c_int64, _P]),
    "vt_mlp_score": (ctypes.c_int, [_P, ctypes.POINTER(Camera), ctypes.c_int32, ctypes.c_int32, _P,
                                    ctypes.c_int64, _P]),
    "vt_file_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_P),
                                    ctypes.POINTER(ctypes.c_int64)]),
    "vt_file_close": (ctypes.c_int, [_P]),
    "vt_file_read_box": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), _P]),
    "vt_file_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _P, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.c_int64, ctypes.c_int32]),
    "vt_volume_write_bricks_hdf5": (ctypes.c_int, [_P, _P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                                   ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P, ctypes.c_int64,
                                                   ctypes.c_char_p, ctypes.c_char_p]),
    "vt_cache_attach_hdf5_bricks": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_char_p]),
    "vt_file_cut": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_int32]),
}


def lib() -> ctypes.CDLL:
    """Load libvt_hip.so (built by __graft_entry__.build()); raises if absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, ctx=None) -> None:
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        lib().vt_last_error(ctx, buf, len(buf))
        raise VTError(rc, buf.value.decode(errors="replace"))


def screen_size(distance: float, fov_h: float, fov_w: float) -> tuple[float, float]:
    """Screen::Screen plane size (Screen.cpp:18-19); fov arguments are half-angles."""
    h, w = ctypes.c_float(), ctypes.c_float()
    _check(lib().vt_screen_size(distance, fov_h, fov_w, ctypes.byref(h), ctypes.byref(w)))
    return h.value, w.value


def look_at(eye, target, up_hint=(0.0, 1.0, 0.0)):
    """Orthonormal (look, up, right) with right = look x up as in Camera.cu:20."""
    e = np.asarray(eye, np.float64)
    t = np.asarray(target, np.float64)
    look = t - e
    look /= np.linalg.norm(look)
    right = np.cross(look, np.asarray(up_hint, np.float64))
    right /= np.linalg.norm(right)
    up = np.cross(right, look)
    return look.astype(np.float32), up.astype(np.float32), right.astype(np.float32)


def write_ppm(path, rgba: np.ndarray, W: int, H: int) -> None:
    """Write a host RGBA frame (W*H x 4 float32) as a PPM laid out like the
    reference's FreeImage bitmap (vt_image_write_ppm)."""
    a = np.ascontiguousarray(rgba, np.float32)
    if a.size != W * H * 4:
        raise ValueError("rgba must hold W*H float4 pixels")
    _check(lib().vt_image_write_ppm(a.ctypes.data, W, H, str(path).encode()))


def make_camera(W: int, H: int, pos, look=(0, 0, -1), up=(0, 1, 0), right=None, fov_v_deg: float = 45.0,
                distance: float = 50.0, fov_w_half_deg: float | None = None) -> Camera:
    """Camera with a *full* vertical field of view fov_v_deg: the reference's
    Screen takes a half-angle (Screen.cpp:18), so fov_v_deg/2 is passed, and
    the horizontal half-angle follows the aspect ratio unless given."""
    look = np.asarray(look, np.float32)
    up = np.asarray(up, np.float32)
    if right is None:
        right = np.cross(look.astype(np.float64), up.astype(np.float64)).astype(np.float32)
    half_v = fov_v_deg / 2.0
    if fov_w_half_deg is None:
        fov_w_half_deg = math.degrees(math.atan(math.tan(math.radians(half_v)) * W / H))
    h, w = screen_size(distance, half_v, fov_w_half_deg)
    c = Camera()
    c.pos[:] = [float(v) for v in np.asarray(pos, np.float32)]
    c.look[:] = [float(v) for v in look]
    c.up[:] = [float(v) for v in up]
    c.right[:] = [float(v) for v in np.asarray(right, np.float32)]
    c.h, c.w, c.distance = h, w, distance
    c.W, c.H = W, H
    return c


class DeviceBuffer:
    """Device allocation owned by a Context."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().vt_malloc(ctx.handle, self.nbytes, ctypes.byref(p)), ctx.handle)
        self.ptr = p.value

    def upload(self, arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        _check(lib().vt_memcpy_h2d(self.ctx.handle, self.ptr, a.ctypes.data, a.nbytes), self.ctx.handle)
        return self

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().vt_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes), self.ctx.handle)
        return out

    def download_at(self, offset: int, dtype, shape) -> np.ndarray:
        """Copy `shape` elements starting `offset` bytes into the buffer to the host."""
        out = np.empty(shape, dtype)
        assert 0 <= offset and offset + out.nbytes <= self.nbytes
        _check(lib().vt_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr + int(offset), out.nbytes),
               self.ctx.handle)
        return out

    def zero(self) -> None:
        _check(lib().vt_memset(self.ctx.handle, self.ptr, 0, self.nbytes), self.ctx.handle)

    def free(self) -> None:
        if self.ptr:
            lib().vt_free(self.ctx.handle, self.ptr)
            self.ptr = None


class Context:
    """One vt_context (device, stream, octree, brick cache)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().vt_context_create(device, ctypes.byref(h)))
        self.handle = h.value
        self._keep = []

    def close(self) -> None:
        if self.handle:
            lib().vt_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, rc):
        _check(rc, self.handle)

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr: np.ndarray) -> DeviceBuffer:
        a = np.ascontiguousarray(arr)
        return DeviceBuffer(self, a.nbytes).upload(a)

    def copy_probe(self, src: DeviceBuffer, dst: DeviceBuffer, nbytes: int) -> None:
        self._c(lib().vt_copy_probe(self.handle, src.ptr, dst.ptr, nbytes))

    def synchronize(self) -> None:
        self._c(lib().vt_synchronize(self.handle))

    # ---- octree
    def octree_upload(self, runs: np.ndarray, sizes, n_levels: int, max_level: int) -> None:
        r = np.ascontiguousarray(runs, np.uint64)
        s = np.ascontiguousarray(np.asarray(sizes)[: max_level + 1], np.int32)
        self._c(lib().vt_octree_upload(self.handle, r.ctypes.data, s.ctypes.data, n_levels, max_level))

    def octree_load_file(self, path: str, max_level: int):
        iso, n = ctypes.c_float(), ctypes.c_int32()
        rd = (ctypes.c_int32 * 3)()
        self._c(lib().vt_octree_load_file(self.handle, str(path).encode(), max_level, ctypes.byref(iso),
                                          ctypes.byref(n), rd))
        return iso.value, n.value, tuple(rd)

    def octree_build(self, d_vol: "DeviceBuffer", dims, iso: float, literal: bool, max_level: int):
        """GPU octree builder; returns (nLevels, [run-list size per level])."""
        n = ctypes.c_int32()
        sizes = np.zeros(32, np.int64)
        self._c(lib().vt_octree_build(self.handle, d_vol.ptr, *[int(d) for d in dims], iso, int(literal), max_level,
                                      ctypes.byref(n), sizes.ctypes.data))
        return n.value, [int(v) for v in sizes[: n.value + 1]]

    def octree_runs(self, level: int, size: int) -> np.ndarray:
        out = np.zeros(max(size, 1), np.uint64)
        self._c(lib().vt_octree_get_runs(self.handle, level, out.ctypes.data, size))
        return out[:size]

    def octree_write_file(self, path) -> None:
        self._c(lib().vt_octree_write_file(self.handle, str(path).encode()))

    def volume_generate(self, kind: int, dims, seed: int, d_out: "DeviceBuffer", shapes=None, n: int = 0,
                        octaves: int = 0, base_freq: float = 0.0, gain: float = 0.5, lacunarity: float = 2.0,
                        scale: float = 1.0, x0: int = 0, nx: int | None = None) -> None:
        X, Y, Z = (int(d) for d in dims)
        p = GenParams()
        p.kind, p.seed, p.n = kind, seed & 0xFFFFFFFF, n
        sh = np.ascontiguousarray(shapes if shapes is not None else np.zeros(0), np.float32)
        stride = 8 if kind == GEN_SDF else 5
        p.count = sh.size // stride if sh.size else 0
        p.shapes = sh.ctypes.data if sh.size else None
        p.octaves, p.base_freq, p.gain, p.lacunarity, p.scale = octaves, base_freq, gain, lacunarity, scale
        nx = X - x0 if nx is None else nx
        self._c(lib().vt_volume_generate(self.handle, ctypes.byref(p), X, Y, Z, x0, nx, d_out.ptr))

    def volume_write_hdf5(self, d_vol: "DeviceBuffer", dims, path, dataset: str = "value", chunk: int = 0) -> None:
        X, Y, Z = (int(d) for d in dims)
        self._c(lib().vt_volume_write_hdf5(self.handle, d_vol.ptr, X, Y, Z, str(path).encode(), dataset.encode(),
                                           chunk))

    # ---- cache
    def cache_create(self, level_c: int, cube_dim: int, cube_inc: int, gpu_slots: int, cpu_slots: int = 0):
        self._c(lib().vt_cache_create(self.handle, level_c, cube_dim, cube_inc, gpu_slots, cpu_slots))

    def cache_destroy(self) -> None:
        self._c(lib().vt_cache_destroy(self.handle))

    def cache_attach_device_volume(self, vol: DeviceBuffer, dims) -> None:
        self._keep.append(vol)
        self._c(lib().vt_cache_attach_device_volume(self.handle, vol.ptr, *[int(d) for d in dims]))

    def cache_attach_hdf5_bricks(self, file: str, dataset: str = "bricks") -> None:
        """Brick source = a brick-major file (volume_write_bricks_hdf5)."""
        self._c(lib().vt_cache_attach_hdf5_bricks(self.handle, str(file).encode(), dataset.encode()))

    def volume_write_bricks_hdf5(self, d_vol: DeviceBuffer, dims, path, level_c: int, cube: int, inc: int, codes,
                                 dataset: str = "bricks") -> None:
        """The bricks `codes` (ascending level-level_c codes) of a device volume
        as a brick-major HDF5 file, one zero-padded brick with halo per record."""
        c = np.ascontiguousarray(codes, np.uint64)
        self._c(lib().vt_volume_write_bricks_hdf5(self.handle, d_vol.ptr, *[int(d) for d in dims], level_c, cube,
                                                  inc, c.ctypes.data, c.size, str(path).encode(), dataset.encode()))

    def cache_attach_host_from_device(self, vol: DeviceBuffer, dims) -> None:
        """A pinned host tier holding the bricks with nodes, filled from a device volume."""
        self._c(lib().vt_cache_attach_host_from_device(self.handle, vol.ptr, *[int(d) for d in dims]))

    def cache_attach_hdf5(self, file: str, dataset: str) -> None:
        self._c(lib().vt_cache_attach_hdf5(self.handle, str(file).encode(), dataset.encode()))

    def cache_prefetch_host(self, wait: bool = True) -> None:
        self._c(lib().vt_cache_prefetch_host(self.handle, 1 if wait else 0))

    def cache_load_bricks(self, codes) -> None:
        c = np.ascontiguousarray(codes, np.uint64)
        self._c(lib().vt_cache_load_bricks(self.handle, c.ctypes.data, c.size))

    def cache_read_brick(self, code: int, S: int) -> np.ndarray:
        """The device slot of a resident brick as an S^3 array (vt_cache_read_brick)."""
        out = np.empty((S, S, S), np.float32)
        self._c(lib().vt_cache_read_brick(self.handle, int(code), out.ctypes.data))
        return out

    def cache_clear(self) -> None:
        self._c(lib().vt_cache_clear(self.handle))

    def cache_stats(self) -> dict:
        s = CacheStats()
        self._c(lib().vt_cache_get_stats(self.handle, ctypes.byref(s)))
        return s.as_dict()

    # ---- render
    def set_frame_options(self, collect_stats: bool = False, record_trace: bool = False) -> None:
        self._c(lib().vt_set_frame_options(self.handle, int(collect_stats), int(record_trace)))

    def render_frame(self, cam: Camera, params: RenderParams, d_rgba: DeviceBuffer, d_hits: DeviceBuffer | None):
        st = FrameStats()
        self._c(lib().vt_render_frame(self.handle, ctypes.byref(cam), ctypes.byref(params), d_rgba.ptr,
                                      d_hits.ptr if d_hits is not None else None, ctypes.byref(st)))
        return st

    def set_march_options(self, skip_empty: bool = True, max_defer: int = 1, persistent: bool = False,
                          max_defer_late: int | None = 2) -> None:
        """Defaults are the library's; max_defer_late=None uses max_defer in every pass."""
        late = max_defer if max_defer_late is None else max_defer_late
        self._c(lib().vt_set_march_options(self.handle, int(skip_empty), int(max_defer), int(late), int(persistent)))

    def set_wide_march(self, lane_budget: int = 1 << 18, first_pass_budget: int = 1 << 21) -> None:
        """Later passes give each parked ray up to 64 lanes that march its
        nodes together, within rays x lanes <= lane_budget (0: off; -k:
        2^k lanes in every later pass); first_pass_budget likewise for the
        first pass."""
        self._c(lib().vt_set_wide_march(self.handle, int(lane_budget), int(first_pass_budget)))

    def set_early_loads(self, reserve: int = 256) -> None:
        """Keep `reserve` free slots before each first pass and copy the
        bricks it requests into them while it runs (0: off)."""
        self._c(lib().vt_set_early_loads(self.handle, int(reserve)))

    def set_first_pass_budget(self, nodes: int = 16) -> None:
        """Rays park after `nodes` walk nodes in the first pass and finish
        in the later passes (0: no limit)."""
        self._c(lib().vt_set_first_pass_budget(self.handle, int(nodes)))

    def set_first_pass_tiles(self, supertile: int = 4) -> None:
        """Deal S x S blocks of 8x8 tiles to the XCDs in the one-lane first
        passes, of single frames and of frames in flight (1: single tiles;
        4: the default)."""
        self._c(lib().vt_set_first_pass_tiles(self.handle, int(supertile)))

    def brick_trace(self, level_c: int) -> np.ndarray:
        n = 1 << (3 * level_c)
        out = np.empty(n, np.uint8)
        self._c(lib().vt_get_brick_trace(self.handle, out.ctypes.data, n))
        return out

    def trace_nodes(self, cam: Camera, final_level: int, max_nodes: int, cap_total: int):
        R = cam.W * cam.H
        nodes = self.alloc(max(R * max_nodes * 8, 8))
        counts = self.alloc(R * 4)
        try:
            self._c(lib().vt_trace_nodes(self.handle, ctypes.byref(cam), final_level, max_nodes, cap_total,
                                         nodes.ptr, counts.ptr))
            n = nodes.download(np.uint64, (R, max_nodes))
            c = counts.download(np.int32, (R,))
        finally:
            nodes.free()
            counts.free()
        return n, c

    def camera_rays(self, cam: Camera) -> np.ndarray:
        R = cam.W * cam.H
        d = self.alloc(R * 12)
        try:
            self._c(lib().vt_camera_rays(self.handle, ctypes.byref(cam), d.ptr))
            return d.download(np.float32, (R, 3))
        finally:
            d.free()


class MLP:
    """The brick-visibility MLP (vt_mlp_*)."""

    def __init__(self, ctx: Context, seed: int = 1, max_batch: int = 8192):
        h = ctypes.c_void_p()
        _check(lib().vt_mlp_create(ctx.handle, seed, max_batch, ctypes.byref(h)), ctx.handle)
        self.ctx, self.handle, self.max_batch = ctx, h.value, max_batch
        self.nparams = lib().vt_mlp_num_params()

    def close(self) -> None:
        if self.handle:
            lib().vt_mlp_destroy(self.handle)
            self.handle = None

    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, np.float32)
        _check(lib().vt_mlp_get_params(self.handle, out.ctypes.data))
        return out

    def set_params(self, p) -> None:
        a = np.ascontiguousarray(p, np.float32)
        assert a.size == self.nparams
        _check(lib().vt_mlp_set_params(self.handle, a.ctypes.data))

    def forward(self, d_x: DeviceBuffer, n: int, d_out: DeviceBuffer) -> None:
        _check(lib().vt_mlp_forward(self.handle, d_x.ptr, n, d_out.ptr))

    def gradients(self, d_x: DeviceBuffer, d_y: DeviceBuffer, n: int, pos_weight: float = 1.0):
        g = np.zeros(self.nparams, np.float32)
        loss = ctypes.c_float()
        _check(lib().vt_mlp_gradients(self.handle, d_x.ptr, d_y.ptr, n, pos_weight, g.ctypes.data,
                                      ctypes.byref(loss)))
        return g, loss.value

    def train_step(self, d_x: DeviceBuffer, d_y: DeviceBuffer, n: int, lr: float, pos_weight: float = 1.0) -> float:
        loss = ctypes.c_float()
        _check(lib().vt_mlp_train_step(self.handle, d_x.ptr, d_y.ptr, n, lr, pos_weight, ctypes.byref(loss)))
        return loss.value

    def prefetch(self, cam: Camera, budget: int, min_prob: float = 0.5) -> int:
        """Load the bricks this model expects `cam` to need (vt_prefetch)."""
        n = ctypes.c_int64()
        _check(lib().vt_prefetch(self.ctx.handle, self.handle, ctypes.byref(cam), budget, min_prob,
                                 ctypes.byref(n)), self.ctx.handle)
        return n.value

    def features(self, cam: Camera, n_levels: int, level_c: int, d_codes: DeviceBuffer, n: int,
                 d_x: DeviceBuffer) -> None:
        _check(lib().vt_mlp_features(self.handle, ctypes.byref(cam), n_levels, level_c, d_codes.ptr, n, d_x.ptr))

    def score(self, cam: Camera, n_levels: int, level_c: int, d_codes: DeviceBuffer, n: int,
              d_logits: DeviceBuffer) -> None:
        """Logits of n bricks for the camera in one fused launch (vt_mlp_score)."""
        _check(lib().vt_mlp_score(self.handle, ctypes.byref(cam), n_levels, level_c, d_codes.ptr, n, d_logits.ptr))


class Frames:
    """Frames in flight through the context's out-of-core brick cache
    (vt_frames_*): submit frames, poll them, optionally with a prefetcher.
    While it exists, the context belongs to it."""

    def __init__(self, ctx: Context, max_frames: int = 2, max_rays: int = 1920 * 1080):
        h = ctypes.c_void_p()
        _check(lib().vt_frames_create(ctx.handle, max_frames, max_rays, ctypes.byref(h)), ctx.handle)
        self.ctx, self.handle = ctx, h.value

    def set_prefetch(self, mlp: "MLP | None", budget: int = 0, min_prob: float = 0.8) -> None:
        _check(lib().vt_frames_set_prefetch(self.handle, mlp.handle if mlp is not None else None, budget, min_prob),
               self.ctx.handle)

    def submit(self, cam: Camera, params: RenderParams, d_rgba: DeviceBuffer, d_hits: DeviceBuffer) -> int:
        t = ctypes.c_int64()
        _check(lib().vt_frames_submit(self.handle, ctypes.byref(cam), ctypes.byref(params), d_rgba.ptr, d_hits.ptr,
                                      ctypes.byref(t)), self.ctx.handle)
        return t.value

    def poll(self, ticket: int, wait: bool = True):
        """(done, FrameStats or None); raises if the frame failed."""
        done, st = ctypes.c_int32(), FrameStats()
        _check(lib().vt_frames_poll(self.handle, ticket, 1 if wait else 0, ctypes.byref(done), ctypes.byref(st)),
               self.ctx.handle)
        return (True, st) if done.value else (False, None)

    def stats(self) -> dict:
        st = FramesStats()
        _check(lib().vt_frames_get_stats(self.handle, ctypes.byref(st)), self.ctx.handle)
        return st.as_dict()

    def close(self) -> None:
        if self.handle:
            lib().vt_frames_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Render-request serving loop (vt_server_*)."""

    def __init__(self, ctx: Context, hbm_budget: int = 1 << 30, max_tiles: int = 16384):
        h = ctypes.c_void_p()
        _check(lib().vt_server_create(ctx.handle, hbm_budget, max_tiles, ctypes.byref(h)), ctx.handle)
        self.ctx, self.handle = ctx, h.value

    def set_lanes(self, lane_budget: int = 1 << 21) -> None:
        """Lanes per ray of a batch: the widest 2^k (k <= 3) within the lane
        budget; 0 one lane, -k a fixed 2^k."""
        _check(lib().vt_server_set_lanes(self.handle, int(lane_budget)), self.ctx.handle)

    def add_scene(self, d_vol: DeviceBuffer, dims, level_o: int, level_c: int, cube_inc: int = 3) -> int:
        sid = ctypes.c_int32()
        _check(lib().vt_server_add_scene(self.handle, d_vol.ptr, *[int(d) for d in dims], level_o, level_c,
                                         cube_inc, ctypes.byref(sid)), self.ctx.handle)
        return sid.value

    def add_scene_range(self, d_vol: DeviceBuffer, dims, level_o: int, level_c: int, iso_lo: float, iso_hi: float,
                        cube_inc: int = 3) -> tuple[int, int]:
        """A scene for isovalues in [iso_lo, iso_hi]; returns (scene id, resident bricks)."""
        sid, nres = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().vt_server_add_scene_range(self.handle, d_vol.ptr, *[int(d) for d in dims], level_o, level_c,
                                               cube_inc, iso_lo, iso_hi, ctypes.byref(sid), ctypes.byref(nres)),
               self.ctx.handle)
        return sid.value, nres.value

    @staticmethod
    def _request(r: Request, q: dict):
        """Fills `r` from the dict `q`; returns the host outputs when q["keep"]."""
        r.scene = q["scene"]
        r.cam = q["cam"]
        r.iso = q["iso"]
        r.light[:] = list(q.get("light", (0.0, 512.0, 0.0)))
        r.color_scale = q.get("color_scale", 256.0)
        r.arrival_s = q.get("arrival_s", 0.0)
        if not q.get("keep"):
            return None
        R = q["cam"].W * q["cam"].H
        rgba, hits = np.zeros((R, 4), np.float32), np.zeros(R, HIT_DTYPE)
        r.h_rgba, r.h_hits = rgba.ctypes.data, hits.ctypes.data
        return rgba, hits

    def start(self, mode: int = 0) -> None:
        """Run the serving worker (online form); mode 1 = one request per launch."""
        self._pending = {}
        _check(lib().vt_server_start(self.handle, mode), self.ctx.handle)

    def submit(self, q: dict) -> int:
        """Queue request `q` (as for run); returns its ticket."""
        r = Request()
        outs = self._request(r, q)
        t = ctypes.c_int64()
        _check(lib().vt_server_submit(self.handle, ctypes.byref(r), ctypes.byref(t)), self.ctx.handle)
        self._pending[t.value] = outs   # keeps the host outputs alive until done
        return t.value

    def poll(self, ticket: int, wait: bool = True):
        """(done, result dict, outputs or None); raises if the request failed."""
        done, res = ctypes.c_int32(), RequestResult()
        rc = lib().vt_server_poll(self.handle, ticket, 1 if wait else 0, ctypes.byref(done), ctypes.byref(res))
        outs = self._pending.pop(ticket, None) if done.value else None
        _check(rc, self.ctx.handle)
        if not done.value:
            return False, None, None
        return True, {"admit_s": res.admit_s, "done_s": res.done_s, "latency_ms": res.latency_ms,
                      "launches": res.launches}, outs

    def stop(self) -> dict:
        st = ServeStats()
        _check(lib().vt_server_stop(self.handle, ctypes.byref(st)), self.ctx.handle)
        return st.as_dict()

    def run(self, requests, mode: int = 0):
        """requests: list of dicts {scene, cam, iso, arrival_s, keep (bool)}.  Returns
        (results as a list of dicts, stats dict, outputs {index: (rgba, hits)})."""
        n = len(requests)
        arr = (Request * max(n, 1))()
        outs = {}
        for k, q in enumerate(requests):
            o = self._request(arr[k], q)
            if o is not None:
                outs[k] = o
        res = (RequestResult * max(n, 1))()
        st = ServeStats()
        _check(lib().vt_server_run(self.handle, arr, n, mode, res, ctypes.byref(st)), self.ctx.handle)
        results = [{"admit_s": res[k].admit_s, "done_s": res[k].done_s, "latency_ms": res[k].latency_ms,
                    "launches": res[k].launches} for k in range(n)]
        return results, st.as_dict(), outs

    def close(self) -> None:
        if self.handle:
            lib().vt_server_destroy(self.handle)
            self.handle = None


def render_params(level_o: int, level_c: int, iso: float, light=(0.0, 512.0, 0.0),
                  color_scale: float = 256.0) -> RenderParams:
    p = RenderParams()
    p.levelO, p.levelC, p.iso = level_o, level_c, iso
    p.light[:] = [float(v) for v in light]
    p.color_scale = color_scale
    return p


def file_create(path: str, dataset: str, vol: np.ndarray, chunk: int = 0) -> None:
    """FileManager::CreateFile (FileManager.cu:169-243), optional chunking."""
    v = np.ascontiguousarray(vol, np.float32)
    X, Y, Z = v.shape
    _check(lib().vt_file_create(str(path).encode(), dataset.encode(), v.ctypes.data, X, Y, Z, chunk))


def file_cut(path: str, dataset: str, out_path: str, x: int, y: int, z: int, dim: int) -> None:
    """The cutFile tool (cutFile.cu:13-21): the dim^3 box at (x, y, z) of
    `dataset`, zero outside it, written to out_path as dataset "value"."""
    _check(lib().vt_file_cut(str(path).encode(), dataset.encode(), str(out_path).encode(), int(x), int(y), int(z),
                             int(dim)))


class File:
    """FileManager (FileManager.cu:11-167): read-only HDF5 dataset access."""

    def __init__(self, path: str, dataset: str):
        h = ctypes.c_void_p()
        dims = (ctypes.c_int64 * 3)()
        _check(lib().vt_file_open(str(path).encode(), dataset.encode(), ctypes.byref(h), dims))
        self.handle = h.value
        self.dims = tuple(int(d) for d in dims)

    def read_box(self, s, e) -> np.ndarray:
        s3 = (ctypes.c_int32 * 3)(*[int(v) for v in s])
        e3 = (ctypes.c_int32 * 3)(*[int(v) for v in e])
        out = np.empty(tuple(abs(int(e[k]) - int(s[k])) for k in range(3)), np.float32)
        _check(lib().vt_file_read_box(self.handle, s3, e3, out.ctypes.data))
        return out

    def close(self) -> None:
        if self.handle:
            lib().vt_file_close(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
