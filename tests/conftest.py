import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) GPU and the built library")
    config.addinivalue_line("markers", "slow: long CPU test")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:  # noqa: BLE001
        return False


@pytest.fixture(scope="session")
def gpu():
    """The GPU tests run on a GPU box only; there they require the HIP library (no fallback)."""
    if not gpu_available():
        pytest.fail("GPU test selected but no HIP device is visible")
    from lightsmith_amd import _lib
    _lib.load()
    return "cuda:0"
