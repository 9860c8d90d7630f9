import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libfd2 kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle as po

    if not os.path.exists(po.LIB_PATH):
        po.build()
    po.lib()
    return po


@pytest.fixture(scope="session")
def golden():
    import json

    here = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(here, "manifest.json")) as f:
        man = json.load(f)
    return here, man


@pytest.fixture(scope="session")
def fd2lib():
    """libfd2 for GPU tests: it must be built and a GPU must be present."""
    from hybridmantis_amd import fd2

    fd2.lib()  # raises if libfd2.so is missing: no fallback
    return fd2
