import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libddm.so kernels)")


@pytest.fixture(scope="session")
def gpu():
    """The HIP device must be present for gpu-marked tests; no silent fallback."""
    import torch
    assert torch.cuda.is_available(), "gpu test needs a HIP device"
    from hackathon_amd.engine import load_library
    load_library()
    return torch.device("cuda:0")
