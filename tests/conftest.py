import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


def _has_gpu():
    try:
        import spnngp_amd
        return spnngp_amd.device_count() > 0
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def nngp():
    import spnngp_amd
    spnngp_amd.load()
    return spnngp_amd


def uniform_coords(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, size=(n, 2))
    return c[np.argsort(c[:, 0], kind="stable")]
