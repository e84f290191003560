import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests through the C-ABI")


@pytest.fixture(scope="session")
def ctx():
    import kripke_amd as ka
    if ka.device_count() < 1:
        pytest.fail("no HIP device visible for a gpu-marked test")
    c = ka.Context(0)
    yield c
    c.close()
