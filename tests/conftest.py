import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

REFERENCE_SRC = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")


def has_reference_sources():
    return os.path.exists(os.path.join(REFERENCE_SRC, "flam3.c"))


@pytest.fixture(scope="session")
def ref():
    import pyref
    if not pyref.available("strict"):
        pytest.skip("reference library not built (make -C oracle)")
    return pyref


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def prod():
    import flam3_amd
    flam3_amd.lib()
    return flam3_amd


@pytest.fixture(scope="session")
def prod_lib():
    """The product library for its host-only entry points (no GPU needed)."""
    import flam3_amd
    return flam3_amd.lib()
