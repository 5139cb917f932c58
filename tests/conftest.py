import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _has_gpu() -> bool:
    try:
        import neuronbridge_amd as nb
        return nb.lib().nbx_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Fails (not skips) when a GPU test runs without a usable GPU or library."""
    import neuronbridge_amd as nb
    n = nb.lib().nbx_device_count()
    assert n > 0, "no HIP device visible to libnbx"
    return 0
