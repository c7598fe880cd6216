import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (libtpcf kernels)")


@pytest.fixture(scope="session")
def tp():
    """The product package (3pcf_amd)."""
    return importlib.import_module("3pcf_amd")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as o

    o.load()
    return o


@pytest.fixture(scope="session")
def ctx(tp):
    c = tp.Context(0)
    yield c
    c.close()
