import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def hip():
    """Loads the HIP library and selects device 0 (GPU tests only)."""
    import torch
    import ves3d_amd
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    ves3d_amd.init(0)
    return ves3d_amd


@pytest.fixture(scope="session")
def orc():
    import oracle_lib
    return oracle_lib.load()
