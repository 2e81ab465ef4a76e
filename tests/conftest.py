import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def has_reference():
    return os.path.isdir(os.path.join(REFERENCE, "res", "Scenes"))
