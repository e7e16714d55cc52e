import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmilc_hip.so on cuda:0)")
    config.addinivalue_line("markers", "slow: longer CPU oracle checks")


@pytest.fixture(scope="session")
def oracle():
    import ks_oracle
    return ks_oracle
