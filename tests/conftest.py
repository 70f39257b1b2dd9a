import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def capi_mod():
    from xyz_amd import build, capi
    if not os.path.exists(build.LIB):
        build.build()
    capi.lib()
    return capi


@pytest.fixture(scope="session")
def device(capi_mod):
    from xyz_amd import hip
    if hip.device_count() < 1:
        pytest.skip("no HIP device")
    return hip
