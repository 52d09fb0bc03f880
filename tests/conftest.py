import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


def gpu_present() -> bool:
    import simplemoc_amd as sm
    return sm.device_count() > 0


@pytest.fixture(scope="session")
def golden_index():
    import json
    with open(os.path.join(GOLDEN, "golden_index.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def libc():
    import ctypes
    return ctypes.CDLL(None)
