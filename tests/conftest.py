import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("fused-fourier-ptycho_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built libfpm_amd.so")


def _emit(msg):
    # names the case on stdout before any GPU work, flushed
    print(msg, flush=True)


@pytest.fixture(autouse=True)
def _announce(request):
    if request.node.get_closest_marker("gpu"):
        _emit(f"\n[gpu-case] {request.node.nodeid}")
    yield
