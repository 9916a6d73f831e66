import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libkspir_amd.so")
    config.addinivalue_line("markers", "slow: larger CPU oracle runs (tens of seconds)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    if not os.path.exists(O.LIB):
        O.build()
    O.set_threads(min(8, os.cpu_count() or 1))
    return O
