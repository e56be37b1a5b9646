import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP engine)")
    config.addinivalue_line("markers", "slow: large inputs")


@pytest.fixture(scope="session")
def engine():
    """One engine context for the whole GPU session (one process on the card)."""
    import cjtest

    eng = cjtest.Engine(cache_decoded=1)
    yield eng
    eng.close()
