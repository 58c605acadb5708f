"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("cl-ops_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def clops():
    import clops as m
    m.lib()  # raises if the HIP library is not built
    return m


@pytest.fixture(scope="session")
def oracle():
    import oracle as m
    m.lib()
    return m


@pytest.fixture(scope="session")
def ctx(clops):
    # Fails (does not skip) without a device: GPU tests must run the HIP path.
    return clops.Context(0)


@pytest.fixture(scope="session")
def queue(clops, ctx):
    return clops.Queue(ctx)
