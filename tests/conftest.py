import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbeta_hip.so)")
    config.addinivalue_line("markers", "slow: large workload")


@pytest.fixture(scope="session")
def hip_ctx_factory():
    """Creates device contexts; fails loudly if the HIP library or GPU is missing."""
    from betanoox_amd import capi
    made = []

    def make():
        c = capi.Context(0)
        made.append(c)
        return c

    yield make
    for c in made:
        c.close()
