import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs on the MI355X box)")


@pytest.fixture(scope="session")
def oracle():
    from tests import _oracle
    return _oracle


@pytest.fixture(scope="session")
def ffs():
    import flashfftstencil_amd
    flashfftstencil_amd.load_library()
    return flashfftstencil_amd


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but torch sees no GPU")
    return torch
