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
