"""GPU static-core archives (otg_tree_serialize): byte-identical to the
hand-derived archives of two tiny trees, and field by field equal to the
oracle's node table for larger point, box, 6-D, forest and empty trees, with
the node-storage alternative Create picks (ot_static_linear_core.h:330-344,
:367-372, :573-578)."""
import numpy as np
import pytest

from helpers import (HAND_BOXES_2D, HAND_POINTS_2D, check_archive_vs_table, gaussian_mixture, hand_archive_box_2d,
                     hand_archive_point_2d, lognormal_boxes)
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _tree(ctx, kind, D, geom, box, M, cap):
    import octree_amd

    cls = octree_amd.StaticOrthoTreePointND(D) if kind == "point" else octree_amd.StaticOrthoTreeBoxND(D)
    t = cls(ctx=ctx)
    t.Create(geom, M or None, box, cap)
    return t


def test_hand_point_archive_bytes(gpu_ctx):
    t = _tree(gpu_ctx, "point", 2, np.array(HAND_POINTS_2D), [0, 0, 4, 4], 2, 1)
    assert t.serialize() == hand_archive_point_2d()


def test_hand_box_archive_bytes(gpu_ctx):
    t = _tree(gpu_ctx, "box", 2, np.array(HAND_BOXES_2D), None, 2, 1)
    assert t.serialize() == hand_archive_box_2d()


@pytest.mark.parametrize("kind,D,n,M,cap", [("point", 3, 20000, 0, 16), ("point", 6, 8000, 5, 8),
                                            ("box", 3, 20000, 8, 32), ("box", 2, 3000, 0, 4)])
def test_archive_vs_oracle(gpu_ctx, kind, D, n, M, cap):
    from octree_amd.archive import read_static_archive

    if kind == "point":
        geom = gaussian_mixture(n, D, 8, 0.0, 100.0, 1.0, 8.0, seed=n + D)
        box = [0.0] * D + [100.0] * D
    else:
        geom = lognormal_boxes(n, D, 0.0, 100.0, np.log(0.3), 0.6, seed=n + D)
        box = None
    o = O.OracleTree(D, kind, geom, box, M, cap)
    oe = o.export()
    t = _tree(gpu_ctx, kind, D, geom, box, M, cap)
    si = t.set_info()
    smin, smax = list(si.space_min[:D]), list(si.space_max[:D])
    assert np.array_equal(np.array(smin + smax), oe["space"])
    for large in (False, True):
        a = read_static_archive(t.serialize(large_dataset=large), D, kind == "box", large_dataset=large)
        check_archive_vs_table(a, oe, kind == "box", cap, smin, smax)
        assert a["storage"] != 0  # 255 or more entities do not fit NodeStorage256


# (D, n, M) -> variant index.  IsFit needs both the entity count and the
# largest node count possible for it, min(full tree of depth M, full levels the
# entities can fill + one node per entity on every deeper level), below the
# storage's index limit 255 / 65535.
@pytest.mark.parametrize("D,n,M,expect", [
    (2, 3, 2, 0),      # <= 5 nodes
    (3, 200, 2, 0),    # depth 2 holds at most 1 + 8 = 9 nodes
    (3, 100, 10, 1),   # 100 entities may need 9 + 8 * 100 nodes
    (3, 300, 2, 1),    # 300 entities do not fit a u8 segment
    (3, 70000, 3, 2),  # nor a u16 one
])
def test_node_storage_choice(gpu_ctx, D, n, M, expect):
    from octree_amd.archive import read_static_archive

    pts = np.random.default_rng(n).uniform(0.0, 10.0, (n, D))
    box = [0.0] * D + [10.0] * D
    t = _tree(gpu_ctx, "point", D, pts, box, M, 4)
    a = read_static_archive(t.serialize(), D, False)
    assert a["storage"] == expect
    oe = O.OracleTree(D, "point", pts, box, M, 4).export()
    check_archive_vs_table(a, oe, False, 4, [0.0] * D, [10.0] * D)


def test_forest_set_archive_equals_single_tree(gpu_ctx):
    import octree_amd

    sets = [gaussian_mixture(n, 3, 4, 0.0, 50.0, 1.0, 5.0, seed=n) for n in (5, 900, 12000)]
    f = octree_amd.StaticOctreePoint.forest(sets, 8, None, np.tile([[0.0] * 3 + [50.0] * 3], (3, 1)), ctx=gpu_ctx)
    for s, geom in enumerate(sets):
        single = _tree(gpu_ctx, "point", 3, geom, [0.0] * 3 + [50.0] * 3, 0, 8)
        assert f.serialize(s) == single.serialize()


def test_empty_and_managed_archives(gpu_ctx):
    from octree_amd.archive import read_static_archive

    t = _tree(gpu_ctx, "point", 3, np.zeros((0, 3)), [0, 0, 0, 8, 8, 8], 3, 4)
    a = read_static_archive(t.serialize(), 3, False)
    assert a["storage"] == 0 and len(a["child_begins"]) == 0 and len(a["depth"]) == 0
    assert np.array_equal(a["node_size"], np.stack([np.full(3, 8.0 * 0.5 ** k) for k in range(5)]))
    boxes = lognormal_boxes(500, 3, 0.0, 10.0, np.log(0.5), 0.4, seed=3)
    t = _tree(gpu_ctx, "box", 3, boxes, None, 0, 8)
    a = read_static_archive(t.serialize(entities=boxes), 3, True, managed=True)
    assert np.array_equal(a["entities"], boxes)
