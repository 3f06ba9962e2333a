import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libeq_field.so")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def eqf():
    """The product binding; on a GPU test run the HIP library must load.

    torch bundles its own HIP runtime under the same soname as ROCm's; the one
    loaded first serves the whole process.  Tests that hand torch device
    buffers to the library need torch's, so torch is initialised first."""
    if gpu_available():
        import torch
        torch.cuda.init()
    from eq_amd import field
    field.load_library()
    return field
