import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def vto():
    from oracle import vto as m
    m.build()
    return m


@pytest.fixture(scope="session")
def vt():
    """The HIP library; GPU tests fail loudly if it is not built."""
    import visualtur_amd as m
    m.lib()
    return m


@pytest.fixture(scope="session")
def ctx(vt):
    c = vt.Context(0)
    yield c
    c.close()


def blobs_volume(n: int, count: int, seed: int, dims=None) -> np.ndarray:
    """Sum of isotropic Gaussian blobs in normalized coordinates x=(i+0.5)/n."""
    rng = np.random.default_rng(seed)
    dims = dims or (n, n, n)
    c = rng.uniform(0.0, 1.0, (count, 3))
    s = rng.uniform(0.03, 0.12, count)
    a = rng.uniform(0.5, 1.0, count)
    ax = [(np.arange(d, dtype=np.float64) + 0.5) / n for d in dims]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    v = np.zeros(dims, np.float64)
    for k in range(count):
        r2 = (X - c[k, 0]) ** 2 + (Y - c[k, 1]) ** 2 + (Z - c[k, 2]) ** 2
        v += a[k] * np.exp(-r2 / (2 * s[k] ** 2))
    return v.astype(np.float32)
