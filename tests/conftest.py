import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_golden(name: str) -> list[dict]:
    with open(GOLDEN / name) as f:
        return [json.loads(line) for line in f if line.strip()]


@pytest.fixture(scope="session")
def golden():
    return load_golden
