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


def pytest_terminal_summary(terminalreporter):
    """Achieved relative-L2 errors of the GPU parity comparisons (the largest per test, in
    increasing order), printed and written as parity_errors.json into the run-output folder at
    the repository root (SOI_PARITY_OUT overrides it)."""
    gu = sys.modules.get("gpu_util")
    rec = gu.recorded() if gu else {}
    if not rec:
        return
    import json
    rows = sorted(((max(v), k, len(v)) for k, v in rec.items()), key=lambda t: t[0])
    terminalreporter.write_line("achieved relative L2 errors (largest per test; each comparison "
                                "in order when there are at most 4):")
    for e, k, n in rows:
        each = "  [" + ", ".join(f"{v:.2e}" for v in rec[k]) + "]" if 1 < n <= 4 else f"  ({n})"
        terminalreporter.write_line(f"  {e:.2e}  {k.split('::', 1)[-1]}{each}")
    out = os.environ.get("SOI_PARITY_OUT", os.path.join(ROOT, "gpurun_out"))
    try:
        os.makedirs(out, exist_ok=True)
        with open(os.path.join(out, "parity_errors.json"), "w") as f:
            json.dump({k: v for e, k, n in rows for v in [rec[k]]}, f, indent=1)
    except OSError:
        pass
