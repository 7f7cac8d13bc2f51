"""Test configuration: the `gpu` marker and shared fixtures.

CPU tests (`-m "not gpu"`) check the oracle against the reference's golden vectors and closed
forms, the host logic, and that the C-ABI library exports its symbols.  GPU tests (`-m gpu`) call
the HIP engine through its C-ABI / host module and compare against the oracle.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); runs the HIP engine")


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc
    orc._load()
    return orc
