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


VIEWER = ROOT / "visualtur_amd" / "_build" / "vt_view"
# KeyDown's bindings (drawPixel.cu:127-176) as the oracle's camera moves
VIEW_KEYS = {"a": ("RotateY", 5.0), "d": ("RotateY", -5.0), "w": ("MoveForward", -5.0), "s": ("MoveForward", 5.0),
             "x": ("RotateX", 5.0), "y": ("RotateX", -5.0), "c": ("StrafeRight", -5.0), "v": ("StrafeRight", 5.0),
             "f": ("MoveUpward", -5.0), "r": ("MoveUpward", 5.0), "m": ("RotateZ", -5.0), "n": ("RotateZ", 5.0)}


def test_headless_viewer_frames_match_oracle(vt, vto, tmp_path):
    """The headless viewer reads drawPixel.cu's prompts, builds its visualTur
    parameters (levelCubes = nLevels - 5, octreeLevel = levelCubes + 3,
    cubeInc 2, 35-degree field of view) and redraws after every key; every
    frame it writes equals the oracle's render for the camera the same key
    sequence gives, and an unbound key redraws the same frame."""
    assert VIEWER.exists(), "build() compiles the viewer"
    D = 64
    vol = blobs_volume(D, 12, seed=79)
    iso = 0.3
    h5 = tmp_path / "v.h5"
    vt.file_create(h5, "value", vol, chunk=16)
    n, runs = vto.build_octree(vol, iso, literal=True)
    octf = tmp_path / "v.oct"
    vto.octree_write(octf, iso, 1 << n, (D, D, D), n, runs)
    W, H = 72, 56
    pos = (D // 2, D // 2, 3 * D)
    keys = list("adwsxycvfrmn") + ["z"]
    out = tmp_path / "frames"
    out.mkdir()
    stdin = f"{W}\n{H}\n{pos[0]}\n{pos[1]}\n{pos[2]}\n" + " ".join(keys) + " q\n"
    r = subprocess.run([str(VIEWER), str(h5), "value", str(octf), "0", str(out), "raw"], input=stdin,
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert len(list(out.glob("frame_*.ppm"))) == len(keys) + 1
    level_c = n - 5
    V = vto.Volume(vol)
    O = vto.Octree(runs, n, level_c + 3)
    prm = vto.params(level_c + 3, level_c, iso)
    oc = vto.OCamera()
    oc.look[:] = [0.0, 0.0, -1.0]
    oc.up[:] = [0.0, 1.0, 0.0]
    oc.right[:] = [1.0, -0.0, 0.0]   # cross(look, up) in float, as in the test above
    vto.camera_move(oc, "Move", pos)
    hits = 0
    prev = None
    for k, key in enumerate([None] + keys):
        if key in VIEW_KEYS:
            vto.camera_move(oc, *VIEW_KEYS[key])
        cam = camera_dict(out / f"frame_{k:04d}.camera")
        want = vto.camera_vectors(oc)
        for f in ("pos", "look", "up", "right"):
            assert np.array_equal(np.array(cam[f], np.float32).view(np.uint32), want[f].view(np.uint32)), (k, f)
        rc, hc, _ = vto.render(V, O, cam, prm)
        rg = np.fromfile(out / f"frame_{k:04d}.rgba", np.float32).reshape(-1, 4)
        nan = np.isnan(rc)
        assert np.array_equal(np.isnan(rg), nan), k
        assert np.array_equal(rg[~nan].view(np.uint32), rc[~nan].view(np.uint32)), (k, key)
        ppm = (out / f"frame_{k:04d}.ppm").read_bytes()
        head = f"P6\n{H} {W}\n255\n".encode()
        assert ppm.startswith(head) and len(ppm) == len(head) + 3 * W * H, k
        if key == "z":   # unbound: the same picture again
            assert ppm == prev
        prev = ppm
        hits += int((hc["node"] != 0).sum())
    assert hits > 0.05 * W * H

