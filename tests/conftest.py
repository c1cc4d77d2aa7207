import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP library")


def _have_gpu():
    try:
        from ecl_wavelet_amd import _lib
        return _lib.device_count() > 0
    except Exception:  # noqa: BLE001 - no GPU or library on this host
        return False


@pytest.fixture(scope="session")
def gpu():
    if not _have_gpu():
        pytest.fail("GPU test selected but no GPU / HIP library available")
    from ecl_wavelet_amd import _lib
    return _lib
