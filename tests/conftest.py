"""Shared pytest setup: import paths for the package binding and the oracle,
and the `gpu` marker (tests that need an MI355X; the CPU suite runs with
-m "not gpu")."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "merian-wrs_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built libwrs_amd.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def wrs():
    """The HIP library; on a GPU run a missing library or device is a failure."""
    import pywrs
    pywrs.lib()
    assert pywrs.device_count() > 0, "no HIP device visible"
    return pywrs
