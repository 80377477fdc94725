import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "vectors.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle():
    from oracle import refcpu
    refcpu.lib()
    return refcpu


@pytest.fixture(scope="session")
def bc():
    """The product library; GPU tests fail loudly if it is not built."""
    import bluecrunch_amd
    bluecrunch_amd.check(bluecrunch_amd.lib().bcg_init(0), "bcg_init")
    return bluecrunch_amd
