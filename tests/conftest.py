import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    # torch first: its HIP runtime must be the one the process initialises
    # (tests that stage device buffers through torch share it with libslx)
    import torch
    torch.cuda.init()
    import sparselizard_amd as sa
    c = sa.Context(0)
    yield c
    c.close()
