import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gpu-heipa_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on cuda:0)")


@pytest.fixture(scope="session")
def ctx():
    import torch
    import heipa
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = heipa.Context(0, 4 << 30)
    yield c
    c.close()
