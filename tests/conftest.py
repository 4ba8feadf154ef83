import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the kernels through the C ABI")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o  # oracle/oracle.py
    o.load()
    return o


@pytest.fixture(scope="session")
def vvr():
    from volview_amd import _abi
    _abi.load()
    import volview_amd
    return volview_amd


@pytest.fixture(scope="session")
def ctx():
    from volview_amd.renderer import Context
    c = Context(0)
    yield c
    c.close()
