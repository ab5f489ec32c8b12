"""GPU parity of the static-tree build: the node table libotg builds on the
device equals the oracle's (ot_static_linear_core.h:523-599 restated) bit for
bit -- depths, entity segments, children, node keys, MinPoint / MBR geometry,
sorted location keys, and entity storage up to the order of equal keys."""
import numpy as np
import pytest

from helpers import (MT19937, check_node_table_rules, compare_tables, gaussian_mixture, issue38_boxes,
                     lognormal_boxes, uniform_points)
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _gpu_tree(ctx, kind, D, geom, box, M, cap):
    import octree_amd

    cls = octree_amd.StaticOrthoTreePointND(D) if kind == "point" else octree_amd.StaticOrthoTreeBoxND(D)
    t = cls(ctx=ctx)
    t.Create(geom, M or None, box, cap)
    return t


def _parity(ctx, kind, D, geom, box, M, cap, rules=False):
    o = O.OracleTree(D, kind, geom, box, M, cap)
    assert o.status == 0
    oe = o.export()
    g = _gpu_tree(ctx, kind, D, geom, box, M, cap)
    ge = g.export()
    assert ge["max_depth"] == oe["max_depth"]
    compare_tables(ge, oe, kind == "box")
    if rules:
        ge2 = dict(ge)
        ge2["sorted_code"] = oe["sorted_code"]
        ge2["sorted_depth"] = oe["sorted_depth"]
        check_node_table_rules(ge2, geom.shape[0], cap, D, kind == "box")
    return ge


@pytest.mark.parametrize("D", [2, 3, 6])
@pytest.mark.parametrize("n,cap,M", [(1, 1, 3), (2, 1, 4), (17, 16, 3), (1000, 4, 6), (20000, 16, 0), (5000, 1, 8)])
def test_points_parity(gpu_ctx, D, n, cap, M):
    if D == 6 and M > 10:
        M = 8
    pts = uniform_points(n, D, seed=n + D)
    if n > 100:
        pts[: n // 20] = pts[n // 20: 2 * (n // 20)]  # exact duplicates
    _parity(gpu_ctx, "point", D, pts, None, M, cap, rules=n <= 5000)


def test_points_explicit_box(gpu_ctx):
    pts = uniform_points(50000, 3, 0.0, 100.0, seed=1)
    _parity(gpu_ctx, "point", 3, pts, [0, 0, 0, 100, 100, 100], 10, 16)


def test_points_clustered_6d(gpu_ctx):
    pts = gaussian_mixture(30000, 6, 32, 0.0, 1.0, 0.01, 0.05, seed=2)
    _parity(gpu_ctx, "point", 6, pts, [0] * 6 + [1] * 6, 8, 32)


def test_points_all_identical(gpu_ctx):
    pts = np.tile([[1.5, 2.5, 3.5]], (100, 1))
    pts = np.concatenate([pts, [[0.0, 0.0, 0.0], [4.0, 4.0, 4.0]]])
    _parity(gpu_ctx, "point", 3, pts, None, 5, 4, rules=True)


def test_points_c1_size(gpu_ctx):
    """1,000,000 points U[0,100)^3, depth 10, leaf capacity 16 (the plain single-tree workload)."""
    pts = uniform_points(1_000_000, 3, 0.0, 100.0, seed=2024)
    _parity(gpu_ctx, "point", 3, pts, [0, 0, 0, 100, 100, 100], 10, 16)


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("n,cap,M,sigma", [(1, 1, 3, 0.6), (50, 2, 6, 2.0), (2000, 4, 6, 1.0), (20000, 32, 8, 0.6),
                                           (5000, 8, 5, 2.5)])
def test_boxes_parity(gpu_ctx, D, n, cap, M, sigma):
    b = lognormal_boxes(n, D, 0.0, 1000.0, np.log(0.5 * (1000.0 / 256)), sigma, seed=n + D)
    _parity(gpu_ctx, "box", D, b, None, M, cap, rules=n <= 5000)


def test_boxes_issue38(gpu_ctx):
    rng = MT19937(0)
    for _ in range(3):
        b = issue38_boxes(rng)
        _parity(gpu_ctx, "box", 3, b, None, 10, 5, rules=True)


def test_boxes_many_identical_large(gpu_ctx):
    """Repeated boxes spanning most of the space: all-non-refinable leaves and
    non-refinable prefixes of internal nodes (ot_static_linear_core.h:452-476)."""
    b = lognormal_boxes(3000, 3, 0.0, 100.0, np.log(0.5), 0.6, seed=3)
    big = np.tile([[10.0, 10.0, 10.0, 90.0, 90.0, 90.0]], (40, 1))
    mid = np.tile([[2.0, 2.0, 2.0, 30.0, 30.0, 30.0]], (7, 1))
    b = np.concatenate([b, big, mid, np.tile([[0.0, 0.0, 0.0, 100.0, 100.0, 100.0]], (3, 1))])
    _parity(gpu_ctx, "box", 3, b, [0, 0, 0, 100, 100, 100], 6, 4, rules=True)


def test_boxes_out_of_space_go_to_root(gpu_ctx):
    b = lognormal_boxes(2000, 3, 0.0, 100.0, np.log(1.0), 0.6, seed=4)
    _parity(gpu_ctx, "box", 3, b, [5, 5, 5, 95, 95, 95], 5, 8, rules=True)


def test_c2_size_boxes(gpu_ctx):
    b = lognormal_boxes(400_000, 3, 0.0, 1000.0, np.log(0.5), 0.6, seed=7)
    _parity(gpu_ctx, "box", 3, b, None, 8, 32)


@pytest.mark.parametrize("depth", [10, 18])
def test_forest_parity_sort_words(gpu_ctx, depth):
    """Forest whose sort runs inside each set: at depth 10 the 30 location bits and the
    16-bit index inside a set share one word; at depth 18 (54 bits) they do not and the
    sort moves (key, index) pairs.  Both tables equal one oracle build per set."""
    import octree_amd

    D = 3
    # six sets: 3 set-id bits push the whole key past the location bits' last byte, so
    # the sort runs inside the sets; more sort tiles than sets
    sizes = [40000, 35000, 41000, 38000, 30000, 36000]
    sets = [gaussian_mixture(n, D, 8, 0.0, 1000.0, 2.0, 20.0, seed=40 + i) for i, n in enumerate(sizes)]
    caps = np.array([4, 16, 8, 1, 32, 2], dtype=np.uint32)
    depths = np.full(len(sizes), depth, dtype=np.uint32)
    boxes = np.tile([[0.0] * D + [1000.0] * D], (len(sizes), 1))
    f = octree_amd.StaticOctreePoint.forest(sets, caps, depths, boxes, ctx=gpu_ctx)
    for s, geom in enumerate(sets):
        o = O.OracleTree(D, "point", geom, boxes[s], depth, int(caps[s]))
        compare_tables(f.export(s), o.export(), False)


@pytest.mark.parametrize("kind", ["point", "box"])
def test_forest_parity(gpu_ctx, kind):
    """Batched build of uneven sets equals one oracle build per set."""
    import octree_amd

    rng = np.random.default_rng(5)
    sizes = [0, 1, 3, 17, 250, 4000, 1200, 90, 15000, 2, 600]
    D = 3
    sets = []
    for i, n in enumerate(sizes):
        if kind == "point":
            sets.append(gaussian_mixture(n, D, 8, 0.0, 1000.0, 2.0, 20.0, seed=i) if n else np.zeros((0, D)))
        else:
            sets.append(lognormal_boxes(n, D, 0.0, 1000.0, np.log(2.0), 0.8, seed=i) if n else np.zeros((0, 2 * D)))
    caps = rng.choice([1, 4, 8, 16], len(sizes)).astype(np.uint32)
    cls = octree_amd.StaticOctreePoint if kind == "point" else octree_amd.StaticOctreeBox
    boxes = np.tile([[0.0] * D + [1000.0] * D], (len(sizes), 1)) if kind == "point" else None
    f = cls.forest(sets, caps, None, boxes, ctx=gpu_ctx)
    for s, geom in enumerate(sets):
        fe = f.export(s)
        if geom.shape[0] == 0:
            assert f.GetNodeCount(s) == 0
            continue
        o = O.OracleTree(D, kind, geom, None if boxes is None else boxes[s], 0, int(caps[s]))
        compare_tables(fe, o.export(), kind == "box")
