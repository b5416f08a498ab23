import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "harvest-cuda_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))
GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def core():
    from harvest_cuda import _core
    return _core


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    return o


@pytest.fixture(scope="session")
def gpu(core):
    if not core.cuda_available(0):
        pytest.fail("GPU test collected but no supported gfx950 GPU is visible")
    return core
