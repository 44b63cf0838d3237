import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long-running (full-size or many-step) case")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def psb():
    import psb_project_amd
    psb_project_amd.lib()  # fails loudly if libpsb_sph.so is missing
    return psb_project_amd


@pytest.fixture(scope="session")
def ctx(psb):
    from psb_project_amd.device import Context, device_count
    if device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")
    c = Context(int(os.environ.get("PSB_DEVICE", "0")))
    yield c
    c.close()
