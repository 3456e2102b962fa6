import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU oracle work")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    from project_amd import _lib
    _lib.load()   # raises if the HIP library is missing: GPU tests never pass without it
    return True


def announce(msg):
    """Names the GPU case before it runs (a fault can end the process without a report)."""
    print("\n[case] " + msg, flush=True)
