"""Seeded synthetic volumes of the benchmark configurations (host, numpy).

Large volumes are generated on the GPU by the library; these numpy forms
serve small test volumes and the 256^3 Gaussian-blob configuration.
"""
from __future__ import annotations

import numpy as np

C1_SEED = 0xC0FFEE


def gaussian_blobs(n: int, count: int, seed: int, dims=None, sigma=(0.03, 0.12), amp=(0.5, 1.0)) -> np.ndarray:
    """Sum of `count` isotropic Gaussian blobs A*exp(-|x-c|^2 / (2 s^2)) in
    normalized coordinates x = (i+0.5)/n, centres uniform in [0,1]^3."""
    rng = np.random.default_rng(seed)
    dims = tuple(dims) if dims is not None else (n, n, n)
    c = rng.uniform(0.0, 1.0, (count, 3))
    s = rng.uniform(sigma[0], sigma[1], count)
    a = rng.uniform(amp[0], amp[1], count)
    ax = [((np.arange(d, dtype=np.float64) + 0.5) / n) for d in dims]
    v = np.zeros(dims, np.float64)
    for k in range(count):
        gx = np.exp(-((ax[0] - c[k, 0]) ** 2) / (2 * s[k] ** 2))
        gy = np.exp(-((ax[1] - c[k, 1]) ** 2) / (2 * s[k] ** 2))
        gz = np.exp(-((ax[2] - c[k, 2]) ** 2) / (2 * s[k] ** 2))
        v += a[k] * gx[:, None, None] * gy[None, :, None] * gz[None, None, :]
    return v.astype(np.float32)


C3_SEED = 0x5EED0003


def sdf_shapes(count: int, seed: int, extent: int, radius=(0.01, 0.08)) -> np.ndarray:
    """Parameters of the seeded union of spheres and tori (GEN_SDF layout:
    type, cx, cy, cz, r1, r2, axis, pad per shape; voxel units).  Half the
    shapes are spheres of radius r1; tori have major radius r1 and minor
    radius r2 = r1 * U[0.15, 0.4] around axis 0/1/2.  r1 = U[radius] * extent,
    centres uniform in the cube."""
    rng = np.random.default_rng(seed)
    out = np.zeros((count, 8), np.float32)
    out[:, 0] = rng.integers(0, 2, count)
    out[:, 1:4] = rng.uniform(0.0, extent, (count, 3))
    out[:, 4] = rng.uniform(radius[0], radius[1], count) * extent
    out[:, 5] = out[:, 4] * rng.uniform(0.15, 0.4, count)
    out[:, 6] = rng.integers(0, 3, count)
    return out


C5_SEED = 0x5EED0005


def fourier_modes(count: int, seed: int, extent: int, cycles=(1.0, 32.0)) -> np.ndarray:
    """Parameters of the seeded turbulence-like field (GEN_FOURIER layout:
    kx, ky, kz in radians per voxel, phase, amplitude per mode).  `count`
    modes with wavenumbers log-uniform in `cycles` (cycles per extent),
    directions uniform on the sphere, phases uniform in [0, 2 pi), and
    amplitudes |k|^(-5/3) normalised to sum to 1, so the field
    sum_m a_m cos(k_m . x + phi_m) lies in [-1, 1]."""
    rng = np.random.default_rng(seed)
    cyc = np.exp(rng.uniform(np.log(cycles[0]), np.log(cycles[1]), count))
    d = rng.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    amp = cyc ** (-5.0 / 3.0)
    out = np.zeros((count, 5), np.float32)
    out[:, :3] = d * (2.0 * np.pi * cyc / extent)[:, None]
    out[:, 3] = rng.uniform(0.0, 2.0 * np.pi, count)
    out[:, 4] = amp / amp.sum()
    return out


def fourier_field(modes: np.ndarray, x, y, z) -> np.ndarray:
    """The GEN_FOURIER field at integer voxel coordinates, in float32 with
    the generator's order of operations: per mode the phase argument
    ((kx*x + ky*y) + kz*z) + phi, each step rounded, then a*cos summed in
    mode order."""
    x, y, z = (np.asarray(c, np.float32) for c in (x, y, z))
    v = np.zeros(np.broadcast(x, y, z).shape, np.float32)
    for m in np.asarray(modes, np.float32):
        arg = ((m[0] * x + m[1] * y) + m[2] * z) + m[3]
        v = v + m[4] * np.cos(arg)
    return v
