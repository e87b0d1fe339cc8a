import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "dpmd-public_amd"
for p in (str(ROOT), str(PKG_DIR)):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running check")


@pytest.fixture(scope="session")
def golden():
    import json

    def load(name):
        return json.loads((GOLDEN / name).read_text())

    return load


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o

    o.lib()
    return o


def pytest_collection_modifyitems(config, items):
    # GPU tests print the case they run before launching, so a box that dies
    # leaves the name of the case in the log.
    os.environ.setdefault("PYTHONUNBUFFERED", "1")
