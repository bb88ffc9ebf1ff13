import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def weights_b():
    from leela_amd import nets
    return nets.weight_set("B")


@pytest.fixture(scope="session")
def weights_a():
    from leela_amd import nets
    return nets.weight_set("A")
