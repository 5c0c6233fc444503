"""Shared fixtures.  Tests marked ``gpu`` need an MI355X; everything else runs
on the CPU (oracle, golden vectors, host logic, ABI surface)."""
from __future__ import annotations

import importlib
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
if os.path.join(ROOT, "oracle") not in sys.path:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def bk():
    """The product package (directory name has a dash, so import by string)."""
    return importlib.import_module("bernasconi-krylov_amd")


@pytest.fixture(scope="session")
def orc():
    import oracle as O  # noqa: E402  (tests/ may use the oracle as the checker)
    O.lib()
    return O


@pytest.fixture(scope="session")
def golden():
    def load(name):
        with open(os.path.join(GOLDEN, name)) as f:
            return json.load(f)
    return load


@pytest.fixture(scope="session")
def engine(bk):
    eng = bk.Engine(0)
    yield eng
    eng.close()
