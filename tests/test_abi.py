"""CPU tests of the C-ABI library: it is built, loads, and exports every symbol
include/otg/otg.h declares (no compute calls: this container has no GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "otg" / "otg.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"OTG_API\s+[\w\s\*]+?\b(otg_\w+)\s*\(", text)))


def test_header_declares_the_entry_points():
    syms = declared_symbols()
    for must in ("otg_tree_build", "otg_forest_build", "otg_knn", "otg_range", "otg_ray_first", "otg_collide",
                 "otg_tree_export", "otg_result_copy"):
        assert must in syms


def test_library_exports_every_declared_symbol():
    from octree_amd import _abi

    assert _abi.LIB_PATH.exists(), "libotg.so not built (python -m octree_amd.build)"
    lib = ctypes.CDLL(str(_abi.LIB_PATH))
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert sorted(_abi.EXPORTED) == declared_symbols()


def test_python_binding_signatures_load():
    from octree_amd import _abi

    L = _abi.lib()
    assert L.otg_status_string(0) == b"ok"
    assert L.otg_status_string(5) == b"point outside the given tree box"


def test_product_has_no_oracle_dependency():
    """The product path (package + HIP sources) never imports or links the oracle."""
    for p in list((ROOT / "octree_amd").rglob("*.py")) + list((ROOT / "octree_amd" / "csrc").glob("*")):
        t = p.read_text(errors="ignore")
        assert "oracle" not in t.replace("oracle/", "").lower() or p.name == "build.py", p
