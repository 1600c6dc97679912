"""Writes the hand-derived known-answer fixtures of tests/golden/.

The reference ships no golden data, so these vectors come from the
definitions, worked out independently of the oracle:
  * Morton codes from the bit-interleave definition (mortonCodeUtil.hpp:125-148);
  * octree files for 4^3 volumes with one hot voxel, whose valid cells
    follow from the 8-corner sign test (CreateOctree.cu:481-501) by hand:
    a cell is valid iff one of its corners is the hot voxel;
  * slab distances for axis-aligned rays (exact in binary floating point).
Run:  python tests/golden/make_golden.py
"""
import json
import struct
from pathlib import Path

HERE = Path(__file__).resolve().parent


def morton(x, y, z, level, n):
    if level == 0:
        return 1
    x, y, z = x >> (n - level), y >> (n - level), z >> (n - level)
    code = 1
    for b in range(level - 1, -1, -1):
        code = (code << 3) | (((x >> b) & 1) << 2) | (((y >> b) & 1) << 1) | ((z >> b) & 1)
    return code


def octree_bytes(iso, dim, real, n, runs_by_level):
    out = struct.pack("<ifi3ii", 919278872, iso, dim, *real, n)
    for l in range(n, -1, -1):
        r = runs_by_level[l]
        out += struct.pack("<i", len(r)) + struct.pack(f"<{len(r)}Q", *r)
    return out


def main():
    kat = {"morton": [], "slab": [], "octree": []}
    for (x, y, z, level, n) in [(0, 0, 0, 3, 3), (7, 7, 7, 3, 3), (1, 2, 3, 3, 3), (5, 0, 2, 2, 3),
                                (1023, 0, 512, 10, 10), (3, 3, 3, 1, 2), (0, 0, 0, 0, 5)]:
        kat["morton"].append({"xyz": [x, y, z], "level": level, "n": n, "code": morton(x, y, z, level, n)})
    # (origin, dir, min, max) -> (hit, tnear, tfar): axis-aligned, exact
    kat["slab"] = [
        {"o": [-5.0, 0.5, 0.5], "d": [1.0, 0.0, 0.0], "mn": [0, 0, 0], "mx": [1, 1, 1], "hit": True, "tn": 5.0, "tf": 6.0},
        {"o": [0.5, 9.0, 0.5], "d": [0.0, -1.0, 0.0], "mn": [0, 2, 0], "mx": [1, 4, 1], "hit": True, "tn": 5.0, "tf": 7.0},
        {"o": [0.5, 0.5, 0.5], "d": [0.0, 0.0, 1.0], "mn": [0, 0, 0], "mx": [1, 1, 1], "hit": True, "tn": 0.0, "tf": 0.5},
        {"o": [3.0, 3.0, -1.0], "d": [0.0, 0.0, 1.0], "mn": [0, 0, 0], "mx": [2, 2, 2], "hit": False},
    ]
    # 4^3 volume, one hot voxel (value 1, else 0), iso 0.5.
    #  hot (1,1,1): cells {0,1}^3 are valid = level-2 codes 64..71; level 1: node 8
    #  hot (3,3,3): cells {2,3}^3 (the x/y/z=3 ones see the zero padding at 4)
    #               = codes 120..127; level 1: node 15, which the reference's
    #               inner loop never visits (CreateOctree.cu:550), so in the
    #               literal file levels 1 and 0 are empty
    cases = [
        ("hot111", (1, 1, 1), False, {0: [1, 1], 1: [8, 8], 2: [64, 71]}),
        ("hot111_literal", (1, 1, 1), True, {0: [1, 1], 1: [8, 8], 2: [64, 71]}),
        ("hot333", (3, 3, 3), False, {0: [1, 1], 1: [15, 15], 2: [120, 127]}),
        ("hot333_literal", (3, 3, 3), True, {0: [], 1: [], 2: [120, 127]}),
    ]
    for name, hot, literal, runs in cases:
        data = octree_bytes(0.5, 4, (4, 4, 4), 2, runs)
        (HERE / f"{name}.oct").write_bytes(data)
        kat["octree"].append({"name": name, "hot": list(hot), "literal": literal, "iso": 0.5,
                              "runs": {str(k): v for k, v in runs.items()}, "file": f"{name}.oct"})
    (HERE / "kat.json").write_text(json.dumps(kat, indent=1) + "\n")


if __name__ == "__main__":
    main()
