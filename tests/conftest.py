import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "spacetime-optics_amd", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) HIP device")


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "ref_vectors.npz"), allow_pickle=False))


@pytest.fixture(scope="session")
def oracle_libm():
    from oracle_ctypes import Oracle
    return Oracle("libm")


@pytest.fixture(scope="session")
def oracle_bhm():
    from oracle_ctypes import Oracle
    return Oracle("bhm")
