import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib("kd")
    return O


@pytest.fixture(scope="session")
def kd():
    # torch ships its own HIP runtime; when a test also hands torch device
    # tensors to the library, torch's runtime must be initialized before the
    # library's (the reverse order leaves torch without a device).
    try:
        import torch

        if torch.cuda.device_count() > 0:
            torch.cuda.init()
    except ImportError:
        pass
    import kdigest_amd

    kdigest_amd.load()
    return kdigest_amd
