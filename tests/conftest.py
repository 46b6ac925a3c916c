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
    from oracle.pyoracle import Oracle
    return Oracle(omp=False)


@pytest.fixture(scope="session")
def oracle_omp():
    from oracle.pyoracle import Oracle
    return Oracle(omp=True)


@pytest.fixture(scope="session")
def gpu():
    """The library, with a visible GPU required (GPU tests fail, not skip, without one)."""
    import libfluid_amd as lf
    n = lf.device_count()
    assert n > 0, "no HIP device visible: GPU tests need an MI355X"
    return lf
