import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REFERENCE = "/root/reference"
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def reference_file(name):
    """Path of a data file of the reference tree, or skip when it is absent."""
    p = os.path.join(REFERENCE, name)
    if not os.path.exists(p):
        pytest.skip("reference data file %s not present" % name)
    return p


@pytest.fixture(scope="session")
def ctx():
    import pulsarsr_amd as P
    return P.default_context()
