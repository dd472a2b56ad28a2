import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm GPU) and the built HIP library")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Fails (not skips) when a gpu-marked test runs without a GPU."""
    import torch
    assert torch.cuda.is_available(), "gpu test needs a ROCm device"
    import geosmie_amd._lib as L
    L.lib()
    return torch.device("cuda", 0)
