import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "usermeso-2.0-mdpd_amd")
for p in (PKG, os.path.join(REPO, "oracle"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def orc():
    import orc as _orc
    _orc.lib()
    return _orc


@pytest.fixture(scope="session")
def engine():
    import mdpd355
    mdpd355.load()
    return mdpd355


def uniform_box(n, L, seed, lo=(0.0, 0.0, 0.0)):
    rng = np.random.default_rng(seed)
    L = np.broadcast_to(np.asarray(L, dtype=np.float64), (3,))
    x = np.asarray(lo) + rng.uniform(0, 1, (n, 3)) * L
    v = rng.normal(0, 1, (n, 3))
    v -= v.mean(axis=0)
    # FP32-representable inputs so the oracle and the device see the same state
    return x.astype(np.float32).astype(np.float64), v.astype(np.float32).astype(np.float64)
