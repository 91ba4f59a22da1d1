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


def fd_richardson(oracle, surface, spot=100.0, strike=100.0, rate=0.05, dividend=0.02, n=2000):
    """American put on START..END by the oracle's Crank-Nicolson/PSOR solver,
    Richardson-extrapolated over n and n/2 space and time steps."""
    coarse = oracle.fd_american(surface, us(START), spot, strike, rate, dividend, us(END), -1, n // 2, n // 2)
    fine = oracle.fd_american(surface, us(START), spot, strike, rate, dividend, us(END), -1, n, n)
    return fine + (fine - coarse) / 3


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
