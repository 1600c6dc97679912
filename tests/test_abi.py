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
