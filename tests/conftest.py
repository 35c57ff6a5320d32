import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    """Loads the HIP library on the GPU; a GPU test without one fails loudly."""
    import moljax_amd
    lib = moljax_amd.load_library(True)
    assert lib.mjx_init() == 0
    return lib
