import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device and the built libtdse_hip.so")


def complex_random_sample(size, rng=None):
    """Uniform [0, 1) real and imaginary parts (tests/conftest.py:4-5 of the reference)."""
    rng = np.random.default_rng() if rng is None else rng
    return rng.random(size) + 1j * rng.random(size)


@pytest.fixture(scope="session")
def device_ctx():
    """One device context shared by the GPU tests of a session."""
    from ionization_amd.engine import DeviceContext

    ctx = DeviceContext(device=0, pool_bytes=2 << 30)
    yield ctx
    ctx.close()


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
