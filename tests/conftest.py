import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")


def beam_diag(size=(1.0, 0.25, 0.25)):
    return float(sum(s * s for s in size) ** 0.5)


@pytest.fixture(scope="session")
def engine():
    import octopusengine_amd as oe

    e = oe.Engine(0)
    yield e
    e.close()
