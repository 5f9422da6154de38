import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libchimera_hip on the device)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def port():
    from oracle.fimera_ref import Fimera
    return Fimera("port")


@pytest.fixture(scope="session")
def reference():
    from oracle.fimera_ref import REF_LIB, Fimera
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (the reference sources are not present)")
    return Fimera("reference")


@pytest.fixture(scope="session")
def gpu():
    import chimera_amd.fimera as f
    return f
