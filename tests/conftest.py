import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfaust_hip.so)")


@pytest.fixture(scope="session")
def gpu():
    """The product library, loaded once; GPU tests fail (not skip) without it."""
    import torch
    from faust_amd import _lib
    assert torch.cuda.is_available(), "GPU test without a visible GPU"
    return _lib.lib()
