import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built engine library")
    config.addinivalue_line("markers", "slow: longer statistical runs")


_gpu_count = None


def gpu_count():
    """Number of visible GPUs as reported by the engine library itself."""
    global _gpu_count
    if _gpu_count is None:
        from cuda_mcmc_amd import _lib
        import ctypes
        L = _lib.load()          # raises if the library was not built
        buf = ctypes.create_string_buffer(4096)
        n = L.list_gpus(buf, len(buf))
        _gpu_count = max(n, 0)
    return _gpu_count


@pytest.fixture(autouse=True)
def _require_gpu(request):
    if request.node.get_closest_marker("gpu") is not None and gpu_count() == 0:
        pytest.skip("no GPU visible to the engine")
