import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library's C ABI)")


def _ensure_built():
    from risq_amd.hci import LIB_PATH
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB):
        pyoracle.build()
    if not os.path.exists(LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB):
        pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def gpu_ctx():
    _ensure_built()
    from risq_amd.hci import HciContext
    ctx = HciContext(0)
    yield ctx
    ctx.close()
