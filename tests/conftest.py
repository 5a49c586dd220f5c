import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU-side check")


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on cuda:0.  No GPU is an error here, not a skip: a GPU
    test that cannot reach the kernels has not passed."""
    import torch

    from hashinator_amd import _lib

    if not torch.cuda.is_available():
        raise RuntimeError("gpu-marked test run without a GPU")
    lib = _lib.load()
    torch.cuda.init()
    return lib
