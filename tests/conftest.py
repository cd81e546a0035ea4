"""Shared test set-up: the ``gpu`` marker, import paths and fixtures."""
import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

GOLDEN = Path(__file__).resolve().parent / "golden" / "fixtures.json"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def fixtures():
    return json.loads(GOLDEN.read_text())["fixtures"]
