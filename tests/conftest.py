import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libnts.so")
    config.addinivalue_line("markers", "slow: long-running check")


@pytest.fixture(scope="session")
def ctx():
    """A library context on device 0.  No device -> skip with the reason;
    a missing library is an error (never a silent pass)."""
    from nanotracer_amd import _lib
    from nanotracer_amd.api import GpuCtx

    _lib.lib()  # raises if libnts.so is not built
    try:
        c = GpuCtx(0)
    except _lib.NtsError as e:
        if "no HIP device" in str(e):
            pytest.skip(f"no GPU: {e}")
        raise
    yield c
    c.close()
