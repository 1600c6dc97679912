"""The float32 generator restatement (oracle/gen.py) against what its
definitions fix independently of that code: the hash in Python integers,
value noise at lattice points and cell centres, the fBm normalisation, and
signed distances at points whose distance is known in closed form."""
import numpy as np

from oracle import gen


def _hash_py(x: int) -> int:
    m = 0xFFFFFFFF
    x &= m
    x ^= x >> 16
    x = (x * 0x7FEB352D) & m
    x ^= x >> 15
    x = (x * 0x846CA68B) & m
    x ^= x >> 16
    return x


def test_hash_matches_python_integers():
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.array([0, 1, 2, 0xFFFFFFFF, 0x80000000], np.uint64),
                         rng.integers(0, 1 << 32, 2000, dtype=np.uint64)])
    got = gen.hash32(xs.astype(np.uint32))
    assert [int(v) for v in got] == [_hash_py(int(v)) for v in xs]


def test_lattice_value_definition_and_range():
    rng = np.random.default_rng(2)
    ijk = rng.integers(0, 5000, (3, 3000))
    v = gen.lattice_value(ijk[0], ijk[1], ijk[2], 77)
    for t in range(20):
        i, j, k = (int(c) for c in ijk[:, t])
        m = 0xFFFFFFFF
        h = _hash_py(((i * 0x8DA6B343) & m) ^ _hash_py(((j * 0xD8163841) & m) ^
                                                       _hash_py(((k * 0xCB1AB31F) & m) ^ 77)))
        expect = np.float32(h >> 8) * (np.float32(2.0) / np.float32(16777215.0)) - np.float32(1.0)
        assert v[t] == expect
    assert v.min() >= -1.0 and v.max() <= 1.0 and v.min() < -0.9 and v.max() > 0.9
    assert not np.array_equal(v, gen.lattice_value(ijk[0], ijk[1], ijk[2], 78))


def test_value_noise_at_lattice_points_and_cell_centres():
    rng = np.random.default_rng(3)
    ijk = rng.integers(0, 300, (3, 500))
    # t = 0: the smoothstep weights vanish and the noise is the lattice value
    at = gen.value_noise(*ijk.astype(np.float32), 5)
    assert np.array_equal(at, gen.lattice_value(*ijk, 5))
    # t = 1/2 on every axis: weights 1/2, the mean of the 8 corners
    mid = gen.value_noise(*(ijk.astype(np.float32) + np.float32(0.5)), 5)
    corners = [gen.lattice_value(ijk[0] + a, ijk[1] + b, ijk[2] + c, 5).astype(np.float64)
               for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    assert np.allclose(mid, np.mean(corners, axis=0), atol=1e-6)
    # t = 1/2 along x only: halfway between the x-neighbours
    hx = gen.value_noise(ijk[0] + np.float32(0.5), *ijk[1:].astype(np.float32), 5)
    expect = 0.5 * (gen.lattice_value(*ijk, 5).astype(np.float64) +
                    gen.lattice_value(ijk[0] + 1, ijk[1], ijk[2], 5))
    assert np.allclose(hx, expect, atol=1e-6)


def test_fbm_octaves_and_normalisation():
    rng = np.random.default_rng(4)
    p = rng.uniform(0, 200, (3, 4000)).astype(np.float32)
    # one octave at base frequency 1 is the noise itself, times scale
    one = gen.fbm(*p, 9, 1, 1.0, 0.5, 2.0, 1.0)
    assert np.array_equal(one, gen.value_noise(*p, 9))
    assert np.allclose(gen.fbm(*p, 9, 1, 1.0, 0.5, 2.0, 0.25), 0.25 * one.astype(np.float64), atol=1e-7)
    # six octaves: sum_o 0.5^o noise_o(p * 2^o / 16) / sum_o 0.5^o, octave o
    # seeded seed + 7919 o (float64 restatement)
    six = gen.fbm(*p, 9, 6, 1.0 / 16.0, 0.5, 2.0, 1.0)
    acc = np.zeros(p.shape[1])
    for o in range(6):
        acc += 0.5 ** o * gen.value_noise(*(p * np.float32(2.0 ** o / 16.0)), 9 + 7919 * o)
    assert np.allclose(six, acc / sum(0.5 ** o for o in range(6)), atol=1e-5)
    assert six.min() >= -1.0 and six.max() <= 1.0


def test_sdf_distances_at_known_points():
    n = 256
    sphere = [0, 100.0, 80.0, 60.0, 20.0, 0.0, 0.0, 0.0]
    # torus about axis 1 (y): ring in the x-z plane, major radius 30, minor 6
    torus = [1, 160.0, 150.0, 140.0, 30.0, 6.0, 1.0, 0.0]
    shapes = np.array([sphere, torus], np.float32)

    def f(x, y, z):
        return gen.sdf(shapes, n, np.float32(x), np.float32(y), np.float32(z), 1, 0, 0.0, 0.5, 2.0, 0.0)

    assert f(100, 80, 60) == np.float32(-20.0) / np.float32(n)          # sphere centre
    assert f(120, 80, 60) == 0.0                                       # on the sphere
    assert abs(f(100, 80, 90) - 10.0 / n) < 1e-7                        # outside, along z
    assert abs(f(190, 150, 140) - (-6.0 / n)) < 1e-7                    # on the ring
    assert abs(f(160, 150, 140) - (30.0 - 6.0) / n) < 1e-7              # torus centre (on its axis)
    assert abs(f(160, 150 + 10, 140 + 30) - (np.hypot(0, 10) - 6.0) / n) < 1e-7   # above the ring
    # the union takes the nearer shape; the fBm term adds scale x noise
    with_noise = gen.sdf(shapes, n, np.float32(100), np.float32(80), np.float32(60), 1, 2, 1 / 8, 0.5, 2.0, 0.02)
    noise = gen.fbm(np.float32(100), np.float32(80), np.float32(60), 1, 2, 1 / 8, 0.5, 2.0, 0.02)
    assert with_noise == np.float32(-20.0) / np.float32(n) + noise
