"""Oracle Morton codes (mortonCodeUtil.hpp:10-181) against brute force."""
import itertools

import pytest


def interleave(x, y, z, bits):
    """Closed form: bit b of x,y,z -> bits 3b+2, 3b+1, 3b."""
    c = 0
    for b in range(bits):
        c |= ((x >> b) & 1) << (3 * b + 2)
        c |= ((y >> b) & 1) << (3 * b + 1)
        c |= ((z >> b) & 1) << (3 * b)
    return c


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_encode_decode_bruteforce(vto, n):
    D = 1 << n
    for level in range(0, n + 1):
        side = 1 << (n - level)
        seen = set()
        for x, y, z in itertools.product(range(D), repeat=3):
            code = vto.encode(x, y, z, level, n)
            # node containing voxel (x,y,z) at `level`: sentinel + level triples
            nx, ny, nz = x >> (n - level), y >> (n - level), z >> (n - level)
            expect = (1 << (3 * level)) | interleave(nx, ny, nz, level)
            assert code == expect
            assert vto.level(code) == level
            assert vto.decode(code, level, n) == (nx * side, ny * side, nz * side)
            seen.add(code)
        assert len(seen) == 8 ** level   # bijection onto the level's codes


def test_level_edges(vto):
    assert vto.level(1) == 0
    assert vto.level(0) == -1
    for c in (2, 3, 7):
        assert vto.level(c) == -1
    for l in range(1, 20):
        assert vto.level(1 << (3 * l)) == l
        assert vto.level((1 << (3 * l + 3)) - 1) == l
    # codes of level >= 20 saturate at 19 (mortonCodeUtil.hpp:23-24)
    assert vto.level(1 << 60) == 19


def test_dilate(vto):
    for x in [0, 1, 2, 3, 5, 1023, (1 << 21) - 1, 0x155555]:
        assert vto.dilate(x) == interleave(0, 0, x, 21)


def test_update_coordinates_matches_decode(vto):
    """updateCoordinates (mortonCodeUtil.hpp:151-181) gives decode(next) for the
    three moves of the DFS: child, sibling, and back up to a shallower node."""
    n = 5
    import random
    rnd = random.Random(7)
    for _ in range(500):
        l = rnd.randint(1, n - 1)
        cur = vto.encode(rnd.randrange(32), rnd.randrange(32), rnd.randrange(32), l, n)
        mb = vto.decode(cur, l, n)
        child = (cur << 3) | rnd.randrange(8)
        assert vto.update_coordinates(n, l, cur, l + 1, child, mb) == vto.decode(child, l + 1, n)
        sib = (cur & ~7) | rnd.randrange(8)
        assert vto.update_coordinates(n, l, cur, l, sib, mb) == vto.decode(sib, l, n)
        up = cur >> 3
        if l - 1 >= 1:
            assert vto.update_coordinates(n, l, cur, l - 1, up, mb) == vto.decode(up, l - 1, n)
