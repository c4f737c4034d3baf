"""Shared test setup: import paths, the `gpu` marker, and small numeric helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("jacobi-svd-cuda_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library on cuda:0)")


@pytest.fixture(scope="session")
def torch_cuda():
    """torch with a visible GPU; GPU tests fail loudly (not skip) when the device is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu-marked test started without a visible GPU")
    return torch
