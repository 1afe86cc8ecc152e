import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mcmurchie-davidson_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

BASIS_DIR = os.path.join(PKG, "mmd", "basis")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libmmd.so")
    config.addinivalue_line("markers", "slow: long-running CPU oracle check")


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc
    orc.lib()
    return orc


@pytest.fixture(scope="session")
def basis_dir():
    return BASIS_DIR
