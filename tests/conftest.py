import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "jurassic-scatter_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))

GOLDEN = ROOT / "tests" / "golden"
EXAMPLES = GOLDEN / "examples"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture
def in_examples(monkeypatch):
    """Run with the reference examples directory layout as working directory
    (control files name tab/boxcar and ../refrac/... relative to it)."""
    monkeypatch.chdir(EXAMPLES)
    return EXAMPLES
