"""The C-ABI library loads and exports every function include/*.h declares;
host-only entry points (HDF5 files, screen size) checked without a GPU."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        names |= set(re.findall(r"^int\s+(vt_\w+)\s*\(", h.read_text(), re.M))
    return sorted(names)


def test_every_declared_symbol_is_exported(vt):
    L = vt.lib()
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_class_library_exports_reference_surface():
    lib = ROOT / "visualtur_amd" / "_build" / "libvisualtur.so"
    out = subprocess.run(["nm", "-DC", "--defined-only", str(lib)], capture_output=True, text=True).stdout
    syms = ["visualTur::visualTur(visualTurParams_t)", "visualTur::updateVisibleCubes(float*)",
            "Camera::Camera(int, int, int, float, float, float)", "Camera::StrafeRight(float)",
            "Octree::Octree(char const*, Camera*, int)",
            "Octree::getBoxIntersected(int, vt_visible_cube*, vt_visible_cube*)",
            "lruCache::updateCache(vt_visible_cube*, int, int)",
            "rayCaster::render(Camera*, int, int, int, vt_visible_cube*, int3, int3, float*)",
            "rayCaster::renderCube(Camera*, float*, int3, int3, int3, float*)",
            "FileManager::readHDF5_Voxel_Array(int3, int3, float*)"]
    for sym in syms:
        sym = sym.replace("int3", "HIP_vector_type<int, 3u>")
        assert sym in out, sym


def test_screen_size_matches_oracle(vt, vto):
    for d, fh, fw in [(50.0, 22.5, 22.5), (1.0, 10.0, 30.0), (123.5, 44.0, 1.0)]:
        assert vt.screen_size(d, fh, fw) == vto.screen_size(d, fh, fw)


def test_file_manager_host_paths(vt, vto, tmp_path):
    rng = np.random.default_rng(8)
    vol = rng.standard_normal((21, 17, 13)).astype(np.float32)
    V = vto.Volume(vol)
    for chunk in (0, 8):
        p = tmp_path / f"f{chunk}.h5"
        vt.file_create(p, "value", vol, chunk=chunk)
        with vt.File(p, "value") as f:
            assert f.dims == vol.shape
            for s, e in [((0, 0, 0), (21, 17, 13)), ((-4, -1, -7), (6, 9, 3)), ((18, 15, 11), (26, 22, 19)),
                         ((30, 0, 0), (33, 3, 3)), ((-9, -9, -9), (-2, 4, 4)), ((3, 4, 5), (3, 9, 9))]:
                assert np.array_equal(f.read_box(s, e), V.read_box(s, e)), (s, e)
    with pytest.raises(vt.VTError) as e:
        vt.File(tmp_path / "nope.h5", "value")
    assert e.value.code == -3
    with pytest.raises(vt.VTError):
        vt.File(tmp_path / "f0.h5", "no_such_dataset")


def test_file_cut_is_the_box_read(vt, vto, tmp_path):
    """cutFile.cu:13-21: the dim^3 box at (x, y, z), zero outside the
    dataset, written as a dim^3 dataset "value"."""
    rng = np.random.default_rng(9)
    vol = rng.standard_normal((19, 23, 17)).astype(np.float32)
    V = vto.Volume(vol)
    src = tmp_path / "src.h5"
    vt.file_create(src, "density", vol, chunk=8)
    for x, y, z, dim in [(0, 0, 0, 8), (5, 7, 3, 10), (12, 18, 10, 9), (-3, -2, -5, 6), (30, 0, 0, 4), (0, 0, 0, 32)]:
        out = tmp_path / f"cut_{x}_{y}_{z}_{dim}.h5"
        vt.file_cut(src, "density", out, x, y, z, dim)
        with vt.File(out, "value") as f:
            assert f.dims == (dim, dim, dim)
            got = f.read_box((0, 0, 0), (dim, dim, dim))
        assert np.array_equal(got, V.read_box((x, y, z), (x + dim, y + dim, z + dim))), (x, y, z, dim)
    with pytest.raises(vt.VTError) as e:
        vt.file_cut(tmp_path / "nope.h5", "density", tmp_path / "o.h5", 0, 0, 0, 4)
    assert e.value.code == -3
    with pytest.raises(vt.VTError):
        vt.file_cut(src, "density", tmp_path / "o.h5", 0, 0, 0, 0)


def test_ppm_writer_layout(vt, tmp_path):
    """vt_image_write_ppm lays the frame out as testVisualTur.cu:67-107 lays
    out its bitmap: H wide, W tall, pixel (x=i, y=j from the bottom) = ray
    i*W + j, channels = value*255 truncated."""
    W, H = 3, 2
    rgba = np.zeros((W * H, 4), np.float32)
    rgba[:, 0] = np.arange(W * H) / 10.0        # ray id in red: id/10 * 255 truncated
    rgba[4, 1] = 0.5                             # 127.5 -> 127
    rgba[5, 2] = 7.0                             # clamped to 255
    rgba[1, 1] = np.nan                          # NaN -> 0
    path = tmp_path / "f.ppm"
    vt.write_ppm(path, rgba, W, H)
    data = path.read_bytes()
    assert data.startswith(b"P6\n2 3\n255\n")
    px = np.frombuffer(data[len(b"P6\n2 3\n255\n"):], np.uint8).reshape(3, 2, 3)   # rows top-down, x, rgb
    for r in range(3):
        for x in range(2):
            ray = x * W + (W - 1 - r)
            assert px[r, x, 0] == int(np.float32(ray / 10.0) * np.float32(255.0)), (r, x)
    assert px[3 - 1 - 1, 1, 1] == 127      # ray 4 = (x=1, y=1) -> row 1
    assert px[3 - 1 - 2, 1, 2] == 255      # ray 5 = (x=1, y=2) -> row 0
    assert px[3 - 1 - 1, 0, 1] == 0        # ray 1 = (x=0, y=1) -> row 1, NaN


def test_headless_viewer_refuses_bad_input(tmp_path):
    """The headless viewer (drawPixel.cu without a window) checks its
    arguments and the display size it reads before touching a device, and a
    missing HDF5 file is an error, not a crash."""
    viewer = ROOT / "visualtur_amd" / "_build" / "vt_view"
    assert viewer.exists(), "build() compiles the viewer"
    r = subprocess.run([str(viewer)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    for stdin in ("", "0\n5\n1\n2\n3\n", "8\n8\n1\n"):
        r = subprocess.run([str(viewer), "a.h5", "value", "a.oct"], input=stdin, capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 2 and "positive width" in r.stderr, stdin
    r = subprocess.run([str(viewer), str(tmp_path / "missing.h5"), "value", "a.oct", "0", str(tmp_path)],
                       input="8\n8\n1\n2\n3\n", capture_output=True, text=True, timeout=60)
    assert r.returncode == 4 and "unable to open" in r.stderr

