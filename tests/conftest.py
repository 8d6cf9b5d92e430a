import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def rh3t():
    return importlib.import_module("3t_rh_amd")


@pytest.fixture(scope="session")
def workloads():
    return importlib.import_module("3t_rh_amd.workloads")


@pytest.fixture(scope="session")
def oracle():
    import oracle_bridge
    return oracle_bridge
