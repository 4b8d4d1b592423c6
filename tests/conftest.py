import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ref(orc):
    """The reference's own compiled CPU classes, when they could be built."""
    L = orc.ref_lib()
    if L is None and os.path.isdir("/root/reference"):
        try:
            orc.build(ref=True)
        except Exception:
            pass
        L = orc.ref_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return L
