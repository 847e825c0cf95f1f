import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("u-cuda_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def cpu_engine():
    import ucuda
    return ucuda.Engine(device=-1)


@pytest.fixture(scope="session")
def gpu_engine():
    import torch
    import ucuda
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return ucuda.Engine(device=0)
