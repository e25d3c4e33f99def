import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def has_gpu():
    try:
        import petalo_amd
        return petalo_amd.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: GPU tests are selected with -m gpu and must find a device."""
    import petalo_amd
    n = petalo_amd.device_count()
    assert n > 0, "no HIP device visible"
    return petalo_amd
