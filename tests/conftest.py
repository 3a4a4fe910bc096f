import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and liblsgpu.so")


@pytest.fixture(scope="session")
def gpu():
    """The GPU tests run only on a box with a gfx950 device; they fail loudly otherwise."""
    from lightsprintsdk_amd import native
    n = native.device_count()
    assert n > 0, "no gfx950 device visible: " + native.last_error()
    return n


@pytest.fixture(scope="session")
def oracle():
    import oracle_py
    if not oracle_py.have_oracle():
        oracle_py.build_oracle(ref=False)
    return oracle_py
