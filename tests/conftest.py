import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running test")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_engine_factory():
    """Engine constructor for GPU tests; fails loudly when the HIP library is
    missing (no CPU fallback exists)."""
    from md_engine_amd._lib import Engine

    def make():
        return Engine(0)

    return make
