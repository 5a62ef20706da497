import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "big: full-size workload (tens of GB of device memory)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def gpu():
    """Device 0 through the library's own HIP runtime (torch is not imported in
    GPU tests: its wheel bundles another HIP runtime that cannot share the device)."""
    from relativistic_solver_hllc_amd import scatter as rs

    if rs.device_count() < 1:
        pytest.fail("GPU test run without a visible GPU")
    return 0
