"""The oracle's Create<SeqExec> (Build<false>, ot_static_linear_core.h:417-519), pinned
without the reference:
  * its partition step equals the C++ library's std::partition (libstdc++'s
    two-pointer scheme, the one the reference's Partitioning::partition,
    detail/partitioning.h:215-241, restates) on random inputs; a stable partition
    -- a plausible mistake -- differs from it;
  * the tree has exactly the nodes of the sorted build (same depth, key, own
    entities, children and subtree entity set per node), which the node-table rule
    checker pins, in BFS numbering with consecutive child ranges."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from helpers import gaussian_mixture, lognormal_boxes, uniform_points
from oracle import oracle as O

SRC = Path(__file__).resolve().parent / "cpp" / "std_partition.cpp"


@pytest.fixture(scope="module")
def std_partition(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no C++ compiler")
    exe = tmp_path_factory.mktemp("stdpart") / "std_partition"
    subprocess.run([cxx, "-std=c++17", "-O1", str(SRC), "-o", str(exe)], check=True)

    def run(codes, ids, mask, flag):
        lines = [f"{len(codes)} {mask} {flag}"] + [f"{int(c)} {int(i)}" for c, i in zip(codes, ids)]
        out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True,
                             check=True).stdout.split()
        return int(out[0]), np.array([int(x) for x in out[1:]], np.uint32)

    return run


def test_partition_is_the_library_partition(std_partition):
    rng = np.random.default_rng(7)
    stable_differs = 0
    for case in range(120):
        n = int(rng.integers(0, 90))
        D = int(rng.integers(2, 7))
        codes = rng.integers(0, 1 << (3 * D), n).astype(np.uint64)
        ids = rng.permutation(n).astype(np.uint32)
        shift = int(rng.integers(0, 3)) * D
        mask = ((1 << D) - 1) << shift
        flag = int(codes[0]) & mask if n else 0
        m_lib, ids_lib = std_partition(codes, ids, mask, flag)
        c_or, ids_or, m_or = O.partition_codes(codes, ids, mask, flag)
        assert m_or == m_lib and np.array_equal(ids_or, ids_lib), case
        # a stable partition keeps both parts in input order: not the scheme the build uses
        keep = (codes & np.uint64(mask)) == np.uint64(flag)
        stable = np.concatenate([ids[keep], ids[~keep]])
        stable_differs += int(not np.array_equal(stable, ids_lib))
    assert stable_differs > 50


def _node_map(t):
    nn = len(t["depth"])
    rend = np.zeros(nn, np.int64)
    for i in reversed(range(nn)):
        a = int(t["ebegin"][i])
        cc = int(t["ccount"][i])
        rend[i] = rend[int(t["cbegin"][i]) + cc - 1] if cc else a + int(t["ecount"][i])
    out = {}
    for i in range(nn):
        a = int(t["ebegin"][i])
        key = (int(t["depth"][i]), int(t["key"][i]))
        assert key not in out
        out[key] = (int(t["ecount"][i]), int(t["ccount"][i]), tuple(sorted(t["ids"][a:rend[i]].tolist())))
    return out, rend


@pytest.mark.parametrize("kind,D,gen,M,cap", [("point", 3, "uniform", 5, 8), ("point", 2, "uniform", 7, 1),
                                              ("point", 6, "cluster", 3, 16), ("point", 3, "dups", 6, 4),
                                              ("box", 3, "boxes", 5, 8), ("box", 2, "boxes", 6, 2)])
def test_seq_tree_has_the_sorted_trees_nodes(kind, D, gen, M, cap):
    rng = np.random.default_rng(D * 10 + cap)
    if gen == "uniform":
        geom = uniform_points(4000, D, seed=D)
    elif gen == "cluster":
        geom = gaussian_mixture(4000, D, 4, 0.0, 1.0, 0.01, 0.05, seed=D)
    elif gen == "dups":
        geom = uniform_points(3000, D, seed=5)
        geom[:400] = geom[rng.integers(400, 3000, 400)]
    else:
        geom = lognormal_boxes(3000, D, 0.0, 100.0, seed=D)
    par = O.OracleTree(D, kind, geom, None, M, cap).export()
    seq = O.OracleTree(D, kind, geom, None, M, cap, seq=True).export()
    pm, _ = _node_map(par)
    sm, rend = _node_map(seq)
    assert pm == sm
    # BFS numbering, children consecutive inside the parent's range after its own entities
    nxt = 1
    for i in range(len(seq["depth"])):
        cc, cb = int(seq["ccount"][i]), int(seq["cbegin"][i])
        if not cc:
            continue
        assert cb == nxt
        nxt += cc
        pos = int(seq["ebegin"][i]) + int(seq["ecount"][i])
        for c in range(cb, cb + cc):
            assert int(seq["ebegin"][c]) == pos
            pos = int(rend[c])
        assert pos == rend[i]
    assert nxt == len(seq["depth"])
    assert sorted(seq["ids"].tolist()) == list(range(len(geom)))
