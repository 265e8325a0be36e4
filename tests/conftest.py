import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libstab_hip.so)")


def golden(*parts):
    return os.path.join(GOLDEN, *parts)


@pytest.fixture(scope="session")
def ts_profile():
    from oracle import stab_oracle as so
    return so.read_profile(golden("TStest", "profile.0"))


@pytest.fixture(scope="session")
def gpu_solver():
    from stab_amd.solver import Solver
    s = Solver(0)
    yield s
    s.close()
