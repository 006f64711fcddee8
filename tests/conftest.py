"""pytest configuration: the `gpu` marker and import paths."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "emerge-engine_amd", "oracle", "tests"):
    p = os.path.join(REPO, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import oracle_py
    oracle_py.load()
    return oracle_py


@pytest.fixture(scope="session", autouse=True)
def _torch_claims_device_first(request):
    """Tests that borrow torch for device buffers need torch's HIP runtime to
    initialise the device before the library's does (the other order leaves
    torch without a device); so a run that selects GPU tests starts torch's
    device first."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        try:
            import torch
        except ImportError:
            torch = None
        if torch is not None and torch.cuda.is_available():
            torch.zeros(1, device="cuda")
    yield
