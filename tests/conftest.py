"""Shared test setup: imports the product package and the CPU oracle.

Tests marked ``gpu`` need a ROCm device and run the HIP library through its C
ABI; everything else runs on the CPU (oracle against golden vectors, host logic,
library loading)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X)")


@pytest.fixture(scope="session")
def shtk():
    import __graft_entry__

    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def oracle():
    import oracle as o

    o.lib()
    return o


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return torch.device("cuda:0")
