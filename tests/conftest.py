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
    """Starts torch's GPU context before the tests (they hand torch device
    buffers and streams to the library; pptk_amd loads torch's HIP runtime,
    so both share one runtime)."""
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
