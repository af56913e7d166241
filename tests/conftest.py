import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def _have_gpu() -> bool:
    try:
        import fastimagingpipe_amd as F
        return F.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def fip():
    """The product package; GPU tests require a visible device."""
    import fastimagingpipe_amd as F
    if F.device_count() <= 0:
        pytest.fail("no GPU visible: -m gpu tests must run on the MI355X box")
    return F


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
