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
