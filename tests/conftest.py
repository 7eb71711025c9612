import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# torch first: libmtlp.so must bind to the HIP runtime torch loads (one
# runtime per process); importing it here also keeps its slow first import
# out of any single test's time limit.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (built from oracle/ if the library is missing)."""
    from oracle import build as obuild
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        obuild.build(force=True)
    return O


@pytest.fixture(scope="session")
def mtlp():
    """The product package; its library must already be built."""
    import mtlphys_amd
    mtlphys_amd.capi.lib()
    return mtlphys_amd
