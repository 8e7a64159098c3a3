import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")


@pytest.fixture(scope="session")
def prim_tables():
    from dengo_amd.networks import primordial_tables
    names, tab = primordial_tables()
    return names, tab


@pytest.fixture(scope="session")
def solver(prim_tables):
    from dengo_amd import runtime
    s = runtime.PrimordialSolver(prim_tables[1])
    yield s
    s.close()
