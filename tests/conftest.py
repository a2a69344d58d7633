import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One library context for the whole GPU session (fails loudly without a GPU)."""
    from amps_amd import _lib

    ctx = _lib.Context(0)
    yield ctx
    ctx.close()
