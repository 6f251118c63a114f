import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU case")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    subprocess.check_call(["make", "-s", "-f", os.path.join(ROOT, "oracle", "Makefile"), lib])
    return lib


@pytest.fixture(scope="session")
def oracle():
    _ensure_oracle()
    from tests import oracle_py
    return oracle_py.Oracle()


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "products.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden_div():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "division.json")) as f:
        return json.load(f)
