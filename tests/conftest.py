import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs through the C ABI")
    config.addinivalue_line("markers", "slow: large CPU cases")


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session", autouse=True)
def _torch_before_library(request):
    """When GPU tests are selected, torch's HIP runtime starts before the library's: tests that hand torch device
    tensors to the library need torch to have found the device first (started after the library, it finds none)."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        _has_gpu()
    yield


@pytest.fixture(scope="session")
def gpu():
    if not _has_gpu():
        pytest.skip("no HIP device")
    return True
