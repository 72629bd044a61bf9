import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through libdcf.so)")
    config.addinivalue_line("markers", "slow: long-running")


GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="session")
def lena():
    import numpy as np
    return dict(np.load(os.path.join(GOLDEN, "lena512.npz")))
