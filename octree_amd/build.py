"""Build libotg.so (gfx950 HIP + C ABI) in-tree.

Usage: python -m octree_amd.build [--jobs N]

Every HIP translation unit is compiled by hipcc for gfx950 with FP contraction
off (FP64 results must equal the reference's host arithmetic bit for bit) and
without packed-FP32 instruction selection, then linked into
octree_amd/libotg.so next to this file.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "libotg.so"
SOURCES = ["otg_primitives.hip", "otg_build.hip", "otg_query.hip", "otg_abi.hip", "otg_server.hip", "otg_noalloc.hip", "otg_tune.hip"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: libotg needs ROCm's hipcc to build")


COMMON_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> Path:
    out = OBJ / (Path(src).stem + ".o")
    cmd = [hipcc(), *COMMON_FLAGS, "-I", str(HERE.parent / "include"), "-c", str(CSRC / src), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return out


def build(jobs: int = 4) -> Path:
    OBJ.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *[str(o) for o in objs], "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    """The CPU checker (test infrastructure): oracle/liboc.so."""
    root = HERE.parent
    res = subprocess.run(["make", "-s", "-C", str(root / "oracle")], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return root / "oracle" / "liboc.so"


def build_dropin_test() -> Path:
    """tests/cpp/dropin_test: the C++20 drop-in header compiled against libotg."""
    root = HERE.parent
    src = root / "tests" / "cpp" / "dropin_test.cpp"
    out = src.with_suffix("")
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-std=c++20", "-O2", "-I", str(root / "include"), str(src), "-L", str(HERE), "-lotg",
           "-Wl,-rpath,$ORIGIN/../../octree_amd", "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"drop-in test build failed:\n{res.stdout}\n{res.stderr}")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    args = ap.parse_args(argv)
    print(build(args.jobs))
    print(build_oracle())
    print(build_dropin_test())
    return 0


if __name__ == "__main__":
    sys.exit(main())
