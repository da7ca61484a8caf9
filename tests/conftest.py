"""Test configuration: the `gpu` marker and shared fixtures.

GPU tests run on an MI355X; they call libcmag_hip.so through its C ABI and
compare with the oracle (oracle/liboracle.so, the CPU checker).  CPU tests
pin the oracle itself and check the library's host-side surface.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and runs HIP kernels")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def device():
    if not gpu_available():
        pytest.skip("no GPU visible")
    import torch
    from cmag_amd import _lib
    _lib.lib()  # the HIP library must load; a missing library is a failure, not a skip
    return torch.device("cuda:0")
