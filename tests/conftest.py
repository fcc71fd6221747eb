import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built library")


def gpu_count():
    try:
        from pptk_amd import capi
        return capi.lib().pptk_device_count()
    except Exception:
        return 0


@pytest.fixture(scope="session")
def gpu():
    if gpu_count() < 1:
        pytest.fail("no GPU visible to libpptk_amd")
    return True
