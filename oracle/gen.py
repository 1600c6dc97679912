"""Float32 restatement of the synthetic volume generators -- test
infrastructure only.

The benchmark configurations define their volumes in words (value-noise fBm
on a seeded integer-hash lattice; a union of spheres and tori as a signed
distance plus fBm); the reference reads user-supplied HDF5 files and has no
generator, so there is no reference line to follow.  These functions state
the definitions in numpy, every operation in float32 and in the order the
library's generator kernel evaluates it (no fused multiply-add), so that the
device volume can be compared with them bit for bit:

  lattice value  L(i, j, k) = (h >> 8) * (2 / (2^24 - 1)) - 1, with
                 h = H(i*A ^ H(j*B ^ H(k*C ^ seed))) and H the xorshift-multiply
                 finalizer below (uint32 arithmetic);
  value noise    trilinear interpolation of the 8 lattice values around p
                 with smoothstep weights u = t*t*(3 - 2t), t = p - floor(p);
  fBm            sum_o gain^o * noise(p * base_freq * lac^o, seed + 7919 o),
                 divided by sum_o gain^o, times scale;
  SDF            min over shapes of |p - c| - r (spheres) or
                 |(|p_a - c_a| - R, p_h - c_h)| - r (tori about axis h),
                 divided by the extent n, plus fBm.

tests/test_oracle_gen.py pins these statements against what the definitions
fix independently of this code (lattice points, cell centres, distances at
known points); the GPU tests compare the device volumes with them.
"""
from __future__ import annotations

import numpy as np

_F = np.float32
_U = np.uint32


def hash32(x: np.ndarray) -> np.ndarray:
    """xorshift-multiply integer finalizer on uint32 (wrapping arithmetic)."""
    x = np.asarray(x, _U).copy()
    x ^= x >> _U(16)
    x *= _U(0x7FEB352D)
    x ^= x >> _U(15)
    x *= _U(0x846CA68B)
    x ^= x >> _U(16)
    return x


def _u32(i) -> np.ndarray:
    return np.asarray(i, np.int64).astype(np.uint64).astype(_U)


def lattice_value(ix, iy, iz, seed) -> np.ndarray:
    """Seeded lattice value in [-1, 1] at integer lattice point (ix, iy, iz)."""
    s = _U(int(seed) & 0xFFFFFFFF)
    h = hash32((_u32(ix) * _U(0x8DA6B343)) ^ hash32((_u32(iy) * _U(0xD8163841)) ^
                                                     hash32((_u32(iz) * _U(0xCB1AB31F)) ^ s)))
    c = _F(2.0) / _F(16777215.0)
    return (h >> _U(8)).astype(_F) * c - _F(1.0)


def value_noise(x, y, z, seed) -> np.ndarray:
    x, y, z = (np.asarray(v, _F) for v in (x, y, z))
    fx, fy, fz = np.floor(x), np.floor(y), np.floor(z)
    ix, iy, iz = fx.astype(np.int64), fy.astype(np.int64), fz.astype(np.int64)
    tx, ty, tz = x - fx, y - fy, z - fz
    ux = tx * tx * (_F(3.0) - _F(2.0) * tx)
    uy = ty * ty * (_F(3.0) - _F(2.0) * ty)
    uz = tz * tz * (_F(3.0) - _F(2.0) * tz)
    c000, c100 = lattice_value(ix, iy, iz, seed), lattice_value(ix + 1, iy, iz, seed)
    c010, c110 = lattice_value(ix, iy + 1, iz, seed), lattice_value(ix + 1, iy + 1, iz, seed)
    c001, c101 = lattice_value(ix, iy, iz + 1, seed), lattice_value(ix + 1, iy, iz + 1, seed)
    c011, c111 = lattice_value(ix, iy + 1, iz + 1, seed), lattice_value(ix + 1, iy + 1, iz + 1, seed)
    a = c000 + (c100 - c000) * ux
    b = c010 + (c110 - c010) * ux
    c = c001 + (c101 - c001) * ux
    d = c011 + (c111 - c011) * ux
    e = a + (b - a) * uy
    f = c + (d - c) * uy
    return e + (f - e) * uz


def fbm(x, y, z, seed, octaves, base_freq, gain, lacunarity, scale) -> np.ndarray:
    """Fractal sum of value noise at voxel coordinates (x, y, z)."""
    x, y, z = (np.asarray(v, _F) for v in (x, y, z))
    s = np.zeros(np.broadcast(x, y, z).shape, _F)
    amp, norm, fr = _F(1.0), _F(0.0), _F(base_freq)
    for o in range(int(octaves)):
        sd = (int(seed) + 7919 * o) & 0xFFFFFFFF
        s = s + amp * value_noise(x * fr, y * fr, z * fr, sd)
        norm = _F(norm + amp)
        amp = _F(amp * _F(gain))
        fr = _F(fr * _F(lacunarity))
    if int(octaves) <= 0:
        return s   # no octave: no fBm term
    return s / norm * _F(scale)


def sdf(shapes, n, x, y, z, seed, octaves, base_freq, gain, lacunarity, scale) -> np.ndarray:
    """Union of spheres (type 0: centre, radius r1) and tori (type 1: centre,
    major radius r1, minor radius r2, axis 0/1/2 = z/y/x) as a signed
    distance in voxels divided by the extent n, plus fBm.  `shapes` rows:
    type, cx, cy, cz, r1, r2, axis, pad."""
    x, y, z = (np.asarray(v, _F) for v in (x, y, z))
    d = np.full(np.broadcast(x, y, z).shape, _F(3.0e38), _F)
    for s in np.asarray(shapes, _F):
        dx, dy, dz = x - s[1], y - s[2], z - s[3]
        if s[0] == _F(0.0):
            dk = np.sqrt(dx * dx + dy * dy + dz * dz) - s[4]
        else:
            if s[6] == _F(1.0):
                a0, a1, h = dx, dz, dy
            elif s[6] == _F(2.0):
                a0, a1, h = dy, dz, dx
            else:
                a0, a1, h = dx, dy, dz
            q = np.sqrt(a0 * a0 + a1 * a1) - s[4]
            dk = np.sqrt(q * q + h * h) - s[5]
        d = np.fmin(d, dk)
    return d / _F(n) + fbm(x, y, z, seed, octaves, base_freq, gain, lacunarity, scale)
