import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def ctx():
    from libuipc_amd.device import Context

    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def c1():
    """16^3 FEM system (config C1), no contacts."""
    from libuipc_amd import synth

    return synth.fem_system(16, 0, seed=1)


@pytest.fixture(scope="session")
def c1_contact():
    """16^3 FEM system with 20k contact stencils (C2's recipe at C1 size)."""
    from libuipc_amd import synth

    return synth.fem_system(16, 20000, seed=7)


def pytest_report_header(config):
    return f"libuipc_amd tests, root {ROOT}"
