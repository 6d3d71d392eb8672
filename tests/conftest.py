import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def pack():
    from chatcarlo_amd.datapack import build_pack
    return build_pack()


@pytest.fixture(scope="session")
def dpack(pack):
    """The pack as the device sees it (float32-rounded tables)."""
    return pack.rounded()


@pytest.fixture(scope="session")
def engine():
    from chatcarlo_amd.transport import default_engine
    return default_engine()
