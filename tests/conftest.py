import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


# The oracle's ACCD loop takes the library's default step cap in every test
# that compares the two (its own default, like the reference, is no cap)
ORACLE_ACCD_CAP = 1 << 20


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and runs the HIP library")
    from oracle import pyoracle

    pyoracle.set_accd_max_steps(ORACLE_ACCD_CAP)


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle

    pyoracle.lib()
    return pyoracle
