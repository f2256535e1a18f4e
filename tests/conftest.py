import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bt-am_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

TABLE_DIR = os.path.join(PKG, "potentials")
GOLDEN = os.path.join(ROOT, "tests", "golden")
RC = 6.2934034034


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle(TABLE_DIR)


@pytest.fixture(scope="session")
def gpu():
    """A HIP context with the reference tables.  Fails (does not skip) when the
    library or the device is missing: GPU tests must run the native path."""
    import dmdgpu
    return dmdgpu.Context(0, TABLE_DIR, RC)
