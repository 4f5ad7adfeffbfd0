import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X (gfx950) GPU")


def gpu_available() -> bool:
    """True when the HIP runtime reports a device (checked through libbghip, not torch)."""
    try:
        from braingrid_amd import bghip as B
        B.Context(0).close()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip_ctx():
    if not gpu_available():
        pytest.skip("no GPU")
    from braingrid_amd import bghip as B
    ctx = B.Context(0)
    yield ctx
    ctx.close()
