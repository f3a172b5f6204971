import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def have_gpu():
    try:
        import torch  # noqa: F401  (device probe only)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def mg():
    import multigrid_amd
    multigrid_amd.lib()
    return multigrid_amd


@pytest.fixture(scope="session")
def orc():
    from tests import oracle
    oracle.olib()
    return oracle
