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
    import octree_amd

    ctx = octree_amd.Context(0)
    yield ctx
    ctx.close()
