import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for sub in (ROOT, ROOT / "oracle", ROOT / "walking-into-holes_amd"):
    if str(sub) not in sys.path:
        sys.path.insert(0, str(sub))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
