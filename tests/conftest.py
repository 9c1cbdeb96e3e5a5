import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU run")


def _gpu_count() -> int:
    try:
        import cloverleaf_amd
        return cloverleaf_amd.lib().clv_device_count()
    except Exception:
        return 0


@pytest.fixture(scope="session")
def gpu():
    n = _gpu_count()
    if n < 1:
        pytest.fail("GPU test selected but no HIP device is visible to libclover_mi355x.so")
    return n
