import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library's C ABI)")
    config.addinivalue_line("markers", "slow: large CPU workloads")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ref():
    from oracle import refprobe
    if not refprobe.available():
        pytest.skip("reference library not built (oracle/_ref)")
    return refprobe
