import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (runs through the C ABI on the GPU)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def golden() -> Path:
    return GOLDEN
