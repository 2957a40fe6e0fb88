import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the product library on it")


@pytest.fixture(scope="session")
def prg():
    import poissonrecon_gpu_amd as m
    return m


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    return oracle
