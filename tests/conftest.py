import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "k-wave-python_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libkwgpu.so")


@pytest.fixture(scope="session")
def kwlib():
    from kwave_amd import _abi

    return _abi.lib()


@pytest.fixture(scope="session", autouse=True)
def _torch_gpu_context_first(request):
    """GPU sessions initialise PyTorch's HIP context before the solver library loads.

    PyTorch carries its own HIP runtime; it finds the GPU only if it initialises
    before the system runtime the solver library uses.
    """
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch

        if torch.cuda.is_available():
            torch.cuda.init()
    yield
