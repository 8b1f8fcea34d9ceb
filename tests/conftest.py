import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def gpu_available() -> bool:
    from graphblas_amd import capi
    return capi.lib().gm_device_count() > 0


@pytest.fixture(scope="session")
def grb():
    from graphblas_amd import capi
    capi.init()
    return capi
