import datetime as dt
import json
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def golden(name):
    return json.loads((GOLDEN / name).read_text())


def us(value):
    from kiyosi_amd.market import to_micros
    return to_micros(value)


START = dt.datetime(2025, 1, 1)
END = dt.datetime(2026, 1, 1)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as module
    module.build()
    return module


@pytest.fixture(scope="session")
def gpu():
    """The HIP context on device 0; fails loudly when the engine or GPU is missing."""
    from kiyosi_amd import _native
    return _native.context(0)
