import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "accelerate-llvm_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def stream():
    import acc_hip

    if acc_hip.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need one")
    s = acc_hip.Stream(0)
    yield s
    s.destroy()
