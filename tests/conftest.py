import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtwx.so on the device)")


@pytest.fixture(scope="session")
def backend():
    from twixtor_amd.backend import HipBackend

    b = HipBackend()
    yield b
    b.close()
