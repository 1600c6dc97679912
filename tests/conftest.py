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
    from visualtur_amd.volumes import gaussian_blobs
    return gaussian_blobs(n, count, seed, dims)
