import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """A gis handle on device 0; fails (does not skip) when no GPU is present."""
    if not _gpu_available():
        pytest.fail("GPU test collected on a machine without a visible GPU")
    from cufoam_amd.gis import Gis
    g = Gis(0)
    yield g
    g.close()
