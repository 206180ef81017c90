import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")


@pytest.fixture(scope="session")
def ctx():
    from pyiid_amd import _lib
    c = _lib.Context(0)
    yield c
    c.close()


@pytest.fixture
def fixed_shapes(monkeypatch):
    """New contexts skip the timing-based autotuner and use the fixed launch
    shapes.  Tests that compare runs of two contexts bit for bit need it: the
    tuners choose by measured time, and two contexts (or a run that starts
    under graph capture, which keeps the default shape) can pick different
    bins per wave, which moves the recurrences' anchors and the rounding."""
    monkeypatch.setenv("PYIID_AUTOTUNE", "0")
