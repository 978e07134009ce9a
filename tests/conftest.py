"""Shared fixtures. GPU tests are marked `gpu`; everything else runs on CPU."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "many-lights-gi_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def ctx():
    import mlgi
    if mlgi.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    c = mlgi.Context(0)
    yield c
    c.close()
