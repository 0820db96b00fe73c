"""Shared test setup: import paths, the `gpu` marker, GPU skip logic."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfluidsim.so)")


@pytest.fixture(scope="session")
def fs():
    """The product library; GPU tests fail loudly if it is missing."""
    import myfluidsimulation_amd as m
    from myfluidsimulation_amd import capi

    capi.lib()
    if capi.device_count() < 1:
        pytest.fail("no HIP device visible to libfluidsim.so")
    return m


@pytest.fixture(scope="session")
def oracle():
    import pyoracle

    pyoracle.lib()
    return pyoracle
