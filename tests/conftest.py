import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers",
                            "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def hip():
    """The HIP library; GPU tests fail (not skip) if it is missing."""
    from mlcp_amd._lib import lib
    return lib()
