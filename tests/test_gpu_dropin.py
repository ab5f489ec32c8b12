"""The C++20 drop-in header (include/otg/orthotree_gpu.hpp) used the way the
reference's user code uses the static trees, on the reference's KAT data."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "dropin_test"


@pytest.mark.gpu
def test_cpp_dropin_header():
    assert BIN.exists(), "tests/cpp/dropin_test not built (python -m octree_amd.build)"
    res = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "DROPIN OK" in res.stdout


def test_cpp_dropin_binary_links_libotg():
    assert BIN.exists(), "tests/cpp/dropin_test not built (python -m octree_amd.build)"
    out = subprocess.run(["ldd", str(BIN)], capture_output=True, text=True).stdout
    assert "libotg.so" in out and "not found" not in out.split("libotg.so", 1)[1].splitlines()[0]
