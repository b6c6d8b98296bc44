import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm GPU) and libscarfhip.so")


def say(*a):
    """Names the case before a GPU step (flushed, so a failing run shows where it was)."""
    print(*a, flush=True)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def cuda():
    if not gpu_available():
        pytest.skip("no ROCm GPU")
    import torch
    from scarf_amd import _lib
    _lib.lib()  # loud failure if the library is missing
    return torch.device("cuda:0")
