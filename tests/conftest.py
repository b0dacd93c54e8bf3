import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


@pytest.fixture(scope="session")
def amd():
    import rebound_amd
    rebound_amd.lib()
    return rebound_amd


@pytest.fixture(scope="session")
def gpu(amd):
    if amd.device_count() <= 0:
        pytest.fail("GPU test collected but no HIP device is visible")
    return amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
