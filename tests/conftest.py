import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libwpbr_hip.so")


@pytest.fixture(scope="session")
def ctx():
    from woodenpbrengine_amd import lib
    c = lib.Context(0)
    yield c
    c.close()
