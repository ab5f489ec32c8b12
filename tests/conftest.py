import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (str(ROOT), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs through libotg's C ABI")


@pytest.fixture(scope="session")
def gpu_ctx():
    # torch ships its own HIP runtime; it must initialise the device before libotg's
    # runtime does, or torch reports no GPU (device tensors are handed to libotg)
    try:
        import torch

        torch.cuda.init()
    except ImportError:
        pass
    import octree_amd

    ctx = octree_amd.Context(0)
    yield ctx
    ctx.close()
