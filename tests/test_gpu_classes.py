"""The reference-compatible C++ classes (visualTur, Camera, Octree, lruCache,
rayCaster, FileManager), driven by tests/cpp/drive_classes.cpp the way the
reference's drivers call them, checked against the CPU oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT, blobs_volume

pytestmark = pytest.mark.gpu

DRIVER = ROOT / "visualtur_amd" / "_build" / "drive_classes"
CAM_DTYPE = np.dtype([("pos", "<f4", 3), ("look", "<f4", 3), ("up", "<f4", 3), ("right", "<f4", 3),
                      ("h", "<f4"), ("w", "<f4"), ("distance", "<f4"), ("W", "<i4"), ("H", "<i4")])


def camera_dict(path):
    c = np.fromfile(path, CAM_DTYPE)[0]
    return {k: (c[k].tolist() if c[k].shape else c[k].item()) for k in CAM_DTYPE.names}


def test_class_surface_matches_oracle(vt, vto, tmp_path):
    assert DRIVER.exists(), "build() compiles the class driver"
    dims = (64, 60, 50)
    vol = blobs_volume(64, 12, seed=77, dims=dims)
    iso = 0.3
    h5 = tmp_path / "v.h5"
    vt.file_create(h5, "value", vol, chunk=16)
    n, runs = vto.build_octree(vol, iso, literal=True)   # the reference builder's file
    octf = tmp_path / "v.oct"
    vto.octree_write(octf, iso, 1 << n, dims, n, runs)
    W, H, level_c, level_o, inc = 80, 64, 2, 3, 3
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([str(DRIVER), str(h5), "value", str(octf), str(out), str(W), str(H), str(level_c),
                        str(level_o), str(inc), "7", "20"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    V = vto.Volume(vol)
    O = vto.Octree(runs, n, level_o)
    prm = vto.params(level_o, level_c, iso)
    # the driver's camera moves (testVisualTur.cu:48-50, then per frame a
    # strafe and a 7-degree turn) equal the oracle's Camera.cu:83-156
    D = 64
    oc = vto.OCamera()
    oc.look[:] = [0.0, 0.0, -1.0]
    oc.up[:] = [0.0, 1.0, 0.0]
    oc.right[:] = [1.0, -0.0, 0.0]   # cross(look, up) in float: the y term is (-1)*0 - 0*1 = -0
    vto.camera_move(oc, "Move", (0.5 * D, 0.5 * D, 0.5 * D + 2.5 * D))
    vto.camera_move(oc, "MoveForward", 1.0)
    for m in range(3):
        got = camera_dict(out / f"camera_{m}.bin")
        want = vto.camera_vectors(oc)
        for k in ("pos", "look", "up", "right"):
            assert np.array_equal(np.array(got[k], np.float32).view(np.uint32), want[k].view(np.uint32)), (m, k)
        vto.camera_move(oc, "Move", (0.0, 0.0, 0.0))
        vto.camera_move(oc, "StrafeRight", 0.5 * D / 16.0)
        vto.camera_move(oc, "RotateY", 7.0)
    hits_total = 0
    for m in range(3):
        cam = camera_dict(out / f"camera_{m}.bin")
        rc, hc, _ = vto.render(V, O, cam, prm)
        for kind in ("fused", "staged"):
            rg = np.fromfile(out / f"{kind}_{m}.bin", np.float32).reshape(-1, 4)
            nan = np.isnan(rc)
            assert np.array_equal(np.isnan(rg), nan)
            assert np.array_equal(rg[~nan].view(np.uint32), rc[~nan].view(np.uint32)), (kind, m)
        hits_total += int((hc["node"] != 0).sum())
    assert hits_total > 0.05 * 3 * W * H
    cam = camera_dict(out / "camera_cube.bin")
    rc = vto.render_cube(V, cam, iso, (0.0, 512.0, 0.0), (0, 0, 0), 64)
    rg = np.fromfile(out / "cube.bin", np.float32).reshape(-1, 4)
    nan = np.isnan(rc)
    assert np.array_equal(np.isnan(rg), nan)
    assert np.allclose(rg[~nan], rc[~nan], atol=1.0 / 255, rtol=0)
    assert (rc[:, 0] != 1.0).mean() > 0.05
