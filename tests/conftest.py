import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)
GOLDEN = os.path.join(HERE, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture
def in_golden_dir(monkeypatch):
    """The reference reads ./solitons.txt from the working directory
    (src/ic/solitons.cxx:32); run from the fixture directory."""
    monkeypatch.chdir(GOLDEN)
    return GOLDEN


@pytest.fixture(scope="session")
def qx():
    import qaxions_amd
    qaxions_amd.lib()
    return qaxions_amd
