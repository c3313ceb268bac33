import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def hip():
    """The C-ABI binding with device 0 initialised (GPU tests only).

    torch (used by some GPU tests for device memory) is imported first: its
    wheel carries its own HIP runtime under the same soname, and loaded first
    it is the one libvkdt_hip.so binds to, so the process holds one runtime."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    from vkdt_amd import hip as H
    H.init(0)
    return H
