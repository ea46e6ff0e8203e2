import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: longer CPU statistical test")
    # Build the CPU oracle (and, where the reference sources are mounted, the
    # reference's own eduPIC under oracle/_ref) before collection: the tests
    # that compare against eduPIC are skipped only where it cannot be built,
    # not because nothing built it yet.
    if os.path.isdir("/root/reference"):
        from oracle import oracle
        oracle.build()


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (ctypes binding), built on first use."""
    from oracle import oracle
    oracle.lib()
    return oracle
