import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import exm_oracle
    exm_oracle.lib()
    return exm_oracle


@pytest.fixture(scope="session")
def hip():
    """the HIP library on device 0; a GPU test fails (does not skip) when the library is missing"""
    from exmarg_amd import _hip
    _hip.load()
    _hip.init(0)
    return _hip
