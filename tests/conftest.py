import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libntprop.so")


def _gpu_node_present() -> bool:
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def engine():
    """A libntprop context on device 0 (GPU tests only).  On a machine with a
    GPU node any failure to create it is an error, not a skip."""
    if not _gpu_node_present():
        pytest.skip("no GPU node on this machine")
    from neutrinotelescopes_amd.propagation import get_engine
    return get_engine(0)


@pytest.fixture(scope="session")
def oracle():
    import _oracle
    return _oracle.load()
