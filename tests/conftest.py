import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rstsr-showcase-hessian_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libhesship.so")


def pytest_collection_modifyitems(config, items):
    # The CPU suite never touches a GPU: a gpu-marked test run without a device is an error.
    pass


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test collected but no GPU is available")
    return torch.device("cuda:0")
