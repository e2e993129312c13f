import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


def fcc_rows(a):
    return [[0.0, a / 2, a / 2], [a / 2, 0.0, a / 2], [a / 2, a / 2, 0.0]]


@pytest.fixture(scope="session")
def kat1():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "kat1.json")) as f:
        return json.load(f)
