import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libspx on cuda:0)")


@pytest.fixture(scope="session")
def ctx():
    """One device context for the whole GPU session (device 0)."""
    from spinsim_amd import spx
    print(f"[gpu] opening spx context on device 0 ({spx.LIB_PATH})", flush=True)
    # Not closed explicitly: every device object keeps a reference to its
    # context, so the context is released after the last of them.
    return spx.Context(0)
