import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# torch ships its own HIP runtime under the system runtime's SONAME (libamdhip64.so.7): whichever is
# loaded first serves the whole process, and torch does not find the device on the system one. Load
# torch before any test loads libnrgcore.so.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")


@pytest.fixture(scope="session")
def ctx():
    import torch
    from nrgljubljana_amd import Context
    assert torch.cuda.is_available(), "GPU tests need a visible gfx950 device"
    c = Context(0)
    yield c
    c.close()
