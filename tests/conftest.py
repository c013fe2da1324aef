import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    """A library context on device 0; GPU tests fail (not skip) when the library is absent."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    from meraxes_amd._lib import Context

    c = Context(0)
    yield c
    c.close()
