import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the product kernels")


@pytest.fixture(scope="session")
def cw():
    import cwbvh_amd

    return cwbvh_amd


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle
