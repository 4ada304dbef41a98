import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)
GOLDEN = os.path.join(HERE, "golden")
if GOLDEN not in sys.path:
    sys.path.insert(0, GOLDEN)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_bind
    return oracle_bind.Oracle()


@pytest.fixture(scope="session")
def oracle_glibc():
    import oracle_bind
    return oracle_bind.Oracle(glibc=True)


@pytest.fixture(scope="session")
def gpu():
    """A device context for the pisco_dev_* verification entry points."""
    import gpu_bind
    g = gpu_bind.Gpu()
    yield g
    g.close()
