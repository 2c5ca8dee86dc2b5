import os
import sys

import pytest

try:  # one HIP runtime per process: load torch's before libsoi_hip.so binds one
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (checker).  Built on first use."""
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def soi():
    """The product library; GPU tests fail loudly when it has not been built."""
    import soi_fft_amd
    soi_fft_amd.lib()
    return soi_fft_amd


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return torch
