import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libsodabayes.so")


def pytest_collection_modifyitems(config, items):
    """PyTorch ships its own HIP runtime; it finds no device when it starts
    after libsodabayes.so has initialised the system one.  A session that
    runs GPU tests therefore starts torch's device first (device tensors feed
    the device-resident genotype paths)."""
    if any(item.get_closest_marker("gpu") for item in items):
        config._sb_gpu_session = True


@pytest.fixture(scope="session", autouse=True)
def _torch_device_first(request):
    if getattr(request.config, "_sb_gpu_session", False):
        import torch
        if torch.cuda.is_available():
            torch.zeros(1, device="cuda")
    yield
