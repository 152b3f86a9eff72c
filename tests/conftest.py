"""Shared fixtures.  GPU tests are marked ``gpu`` and run only on an MI355X box;
everything else runs on the CPU (oracle, host logic, ABI surface)."""
from __future__ import annotations

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def mcr():
    import __graft_entry__ as ge

    return ge.load_package()


@pytest.fixture(scope="session")
def orc():
    import __graft_entry__ as ge

    o = ge.load_oracle()
    o.lib()
    return o
