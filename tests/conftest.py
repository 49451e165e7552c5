import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "simferrofluid-cuda_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_sessionstart(session):
    # PyTorch ships its own HIP runtime next to the one libsfp links; it only
    # finds the GPU when it initialises first, so tests that hold device
    # arrays in torch tensors need this before any libsfp call.
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:
        pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (runs on the MI355X box)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o
