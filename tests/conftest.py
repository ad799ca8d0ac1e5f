import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE_DIR = os.path.join(ROOT, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def T():
    return importlib.import_module("titan_c-_amd")


@pytest.fixture(scope="session")
def C():
    return importlib.import_module("titan_c-_amd.configs")


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.lib()
    return oracle
