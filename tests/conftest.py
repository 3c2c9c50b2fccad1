import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU-side checks")


@pytest.fixture
def ref_build():
    """The reference's own translation units built as a checker (oracle/_ref), loaded when a
    test runs (never at collection, so a run that deselects these tests loads nothing);
    skips where the build is absent."""
    import oracle as O
    R = O.ref_lib()
    if R is None:
        pytest.skip("reference library build absent (no /root/reference)")
    return R
