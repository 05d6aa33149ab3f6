import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def xlib():
    from xped_amd import _lib

    return _lib.lib()


@pytest.fixture(scope="session")
def rt_struct(xlib):
    """Structure-only runtime: integer bookkeeping without a GPU."""
    from xped_amd.tensor import Runtime

    return Runtime(-1)


@pytest.fixture(scope="session")
def rt(xlib):
    """GPU runtime.  Fails (does not skip) when no GPU is usable: gpu tests must run natively."""
    from xped_amd.tensor import Runtime

    return Runtime(0)
