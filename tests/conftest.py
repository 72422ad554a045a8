import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs through the HIP library)")


def pytest_collection_modifyitems(config, items):
    # GPU tests must not silently pass on a host without a device.
    pass


@pytest.fixture(scope="session")
def knot_tris():
    from c2a_amd import workloads as W
    tris, idx = W.torus_knot()
    return tris.reshape(-1, 9).copy(), idx
