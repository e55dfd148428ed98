"""Test configuration: the ``gpu`` marker and shared helpers.

GPU tests run on an MI355X and call the engine through its C ABI; they are
skipped automatically where no AMD GPU device node exists.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _have_gpu():
    return os.path.exists("/dev/kfd") or bool(os.environ.get("CC_FORCE_GPU_TESTS"))


def pytest_collection_modifyitems(config, items):
    if _have_gpu():
        return
    skip = pytest.mark.skip(reason="no AMD GPU device on this machine")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture
def gpu_cell():
    """Factory for engine cells that are always released (singleton ABI)."""
    from coldcore_amd.gpu import GpuBackend
    made = []

    def make(axes, R, problem="hamming", mu=1, **kw):
        g = GpuBackend(axes, R, problem=problem, mu=mu, **kw)
        made.append(g)
        return g

    yield make
    for g in made:
        g.close()
