"""Shared test set-up: import paths, the `gpu` marker and library fixtures."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "incidence-system-enumeration_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 device (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import _oracle
    return _oracle.load()


@pytest.fixture(scope="session")
def gpu_ctx():
    import ise_amd
    ctx = ise_amd.Context(capacity=1 << 18, stride=32)
    yield ctx
    ctx.close()
