import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long-running check")


@pytest.fixture(scope="session")
def mp_isdf():
    """The compiled extension; GPU tests fail loudly when it is missing."""
    import mp_isdf_amd
    module = mp_isdf_amd.mp_isdf
    if not module.device_enabled:
        pytest.fail("the mp_isdf extension found no HIP device")
    return module
