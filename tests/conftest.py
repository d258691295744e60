import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: long CPU oracle run")


def gpu_available():
    try:
        import torch  # noqa: F401  (device counting does not initialise HIP on this image)

        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip():
    """The HIP library; GPU tests fail loudly if it or the device is missing."""
    from elastic_amd import lib

    lib.load()
    if not gpu_available():
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X box")
    return lib
