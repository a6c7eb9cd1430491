import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def ctx():
    # torch ships its own HIP runtime: it must initialise the device before the
    # library does, or GPU tests that use torch tensors later find no device
    import torch
    torch.cuda.init()
    from nectarrender_amd import native
    c = native.Context(0)
    yield c
    c.close()
