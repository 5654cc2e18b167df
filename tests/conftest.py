import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built liblm.so")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def gpu():
    """The HIP path; GPU tests fail (not skip) when it is missing."""
    import torch
    from lagrangian_microbes_amd import _lib
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    _lib.load()
    return torch.device("cuda")
