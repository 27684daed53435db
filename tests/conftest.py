import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def lib():
    from mf_psipm_amd import _lib
    return _lib.load()


@pytest.fixture(scope="session")
def gpu(lib):
    from mf_psipm_amd import _lib
    if _lib.device_count() < 1:
        pytest.fail("GPU test started without a visible GPU")
    return lib
