import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))
# hipRTC builds of generated surface programs keep their code-object cache
# inside the build tree, not in the user's home
os.environ.setdefault("AMD_COMGR_CACHE_DIR", str(ROOT / "_build" / "comgr_cache"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP render path)")
    config.addinivalue_line("markers", "slow: long CPU oracle run")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(TESTS / "golden" / "reference_vectors.json") as f:
        return json.load(f)
