"""Run-list membership: both windowed searches of the reference
(Octree.cu:78-142,375-532 and CreateOctree.cu:182-284) against plain set
membership on random well-formed run lists."""
import numpy as np
import pytest


def random_runs(rng, level, density, max_gap=6):
    lo, hi = 1 << (3 * level), (2 << (3 * level)) - 1
    bits = rng.random(hi - lo + 1) < density
    runs, c, N = [], 0, bits.size
    while c < N:
        if bits[c]:
            s = c
            while c + 1 < N and bits[c + 1]:
                c += 1
            runs += [lo + s, lo + c]
        c += 1
    return np.array(runs, np.uint64), bits


@pytest.mark.parametrize("level", [1, 2, 3])
@pytest.mark.parametrize("density", [0.02, 0.2, 0.5, 0.9])
def test_windowed_search_is_exact(vto, level, density):
    rng = np.random.default_rng(level * 100 + int(density * 100))
    for trial in range(6):
        runs, bits = random_runs(rng, level, density)
        if runs.size == 0:
            continue
        lo = 1 << (3 * level)
        for parent in range(1 << (3 * (level - 1)), 2 << (3 * (level - 1))):
            exact = 0
            for k in range(8):
                c = (parent << 3) + k
                if bits[c - lo]:
                    exact |= 1 << k
                assert vto.member_exact(runs, c) == bool(bits[c - lo])
            assert vto.children_mask_device(runs, parent) == exact, (parent, runs[:8])
            assert vto.children_mask_builder(runs, parent) == exact, (parent, runs[:8])


def test_adjacent_and_single_runs(vto):
    # adjacent (unmerged) runs, one-element runs, a single pair (size==2 path)
    cases = [
        [8, 8], [8, 15], [9, 9, 10, 10, 12, 14], [8, 9, 10, 11, 12, 13, 14, 15],
        [64, 70, 72, 72, 127, 127], [100, 100],
    ]
    for runs in cases:
        r = np.array(runs, np.uint64)
        level = 1 if runs[0] < 64 else 2
        lo = 1 << (3 * level)
        for parent in range(lo >> 3, (lo << 1) >> 3):
            exact = sum(1 << k for k in range(8) if vto.member_exact(r, (parent << 3) + k))
            assert vto.children_mask_device(r, parent) == exact
            assert vto.children_mask_builder(r, parent) == exact


def test_empty_level_has_no_members(vto):
    assert vto.children_mask_device(np.zeros(0, np.uint64), 1) == 0
