import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "second-candidate_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "ref: needs oracle/_ref (reference sources compiled here)")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context on device 0 for the whole GPU session."""
    import sc_amd
    c = sc_amd.Ctx(0)
    yield c
    c.close()
