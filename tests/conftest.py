import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# the GPU tests run with the hardware queue count bench.py measures with (16; HIP's default is 4), so that the
# library's worker streams really run side by side; set before any HIP runtime starts.  DCF_TEST_HW_QUEUES runs
# them with another count (4: the library's defaults for a process that keeps HIP's)
os.environ["GPU_MAX_HW_QUEUES"] = os.environ.get("DCF_TEST_HW_QUEUES", "16")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through libdcf.so)")
    config.addinivalue_line("markers", "slow: long-running")


GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="session")
def lena():
    import numpy as np
    return dict(np.load(os.path.join(GOLDEN, "lena512.npz")))


@pytest.fixture(scope="session", autouse=True)
def _torch_runtime_first(request):
    """PyTorch brings its own HIP runtime and must initialise it before libdcf.so loads one (a
    runtime loaded first hides the GPU from torch); the GPU tests that hand torch tensors to the
    library rely on this, whatever order the test files run in."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.init()
        except ImportError:
            pass
    yield
