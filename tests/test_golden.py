"""Oracle (and, on a GPU, the HIP builder) against the hand-derived
known-answer fixtures of tests/golden (see make_golden.py)."""
import json
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"
KAT = json.loads((GOLD / "kat.json").read_text())


def hot_volume(hot):
    v = np.zeros((4, 4, 4), np.float32)
    v[tuple(hot)] = 1.0
    return v


def test_morton_kats(vto):
    for k in KAT["morton"]:
        x, y, z = k["xyz"]
        assert vto.encode(x, y, z, k["level"], k["n"]) == k["code"]


def test_slab_kats(vto):
    for k in KAT["slab"]:
        hit, tn, tf = vto.slab(k["o"], k["d"], k["mn"], k["mx"])
        assert hit == k["hit"]
        if hit:
            assert (tn, tf) == (k["tn"], k["tf"])


@pytest.mark.parametrize("case", KAT["octree"], ids=lambda c: c["name"])
def test_octree_builder_kats(vto, tmp_path, case):
    n, runs = vto.build_octree(hot_volume(case["hot"]), case["iso"], literal=case["literal"])
    assert n == 2
    for l in range(3):
        assert runs[l].tolist() == case["runs"][str(l)]
    out = tmp_path / "o.oct"
    vto.octree_write(out, case["iso"], 4, (4, 4, 4), n, runs)
    assert out.read_bytes() == (GOLD / case["file"]).read_bytes()


@pytest.mark.gpu
@pytest.mark.parametrize("case", KAT["octree"], ids=lambda c: c["name"])
def test_gpu_builder_kats(vt, ctx, tmp_path, case):
    d = ctx.to_device(hot_volume(case["hot"]))
    n, sizes = ctx.octree_build(d, (4, 4, 4), case["iso"], case["literal"], 2)
    out = tmp_path / "g.oct"
    ctx.octree_write_file(out)
    assert out.read_bytes() == (GOLD / case["file"]).read_bytes()
    d.free()
