import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libkhx.so")
    config.addinivalue_line("markers", "slow: longer CPU oracle checks")


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def gpu():
    """The HIP device; GPU tests fail (not skip) when it or libkhx is missing."""
    import torch
    from khepri_amd import _lib
    _lib.load()
    assert torch.cuda.is_available(), "GPU test run without a HIP device"
    return torch.device("cuda", 0)
