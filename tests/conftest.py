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


def _init_torch_first():
    """Tests that hand torch device buffers to the library need torch's HIP
    runtime initialised before the library's own: torch cannot start its
    runtime once the library's has claimed the device in this process."""
    try:
        import torch
    except Exception:
        return
    torch.cuda.is_available()


@pytest.fixture(scope="session")
def gpu():
    _init_torch_first()
    if gpu_count() < 1:
        pytest.fail("no GPU visible to libpptk_amd")
    return True
