"""Loading static-core archives into GPU trees (otg_tree_deserialize): the reading
side of serialize(BinaryInputArchive&, StaticLinearOrthoTreeCore&)
(serialization.h:359-386).  A loaded tree must re-serialize to the same bytes,
export the archived table, and answer queries as the tree that wrote it; the
reference's own round trip (tests/unittests/serialize.tests.cpp:133-146: a
StaticOctreePointM of two points, then RangeSearch of the tree box) is included,
as are the hand-derived archives; malformed archives fail with INVALID_ARG."""
import struct

import numpy as np
import pytest

from helpers import (HAND_BOXES_2D, HAND_POINTS_2D, gaussian_mixture, hand_archive_box_2d, hand_archive_point_2d,
                     lognormal_boxes, uniform_points, unit_dirs)

pytestmark = pytest.mark.gpu


def _cls(kind, D):
    import octree_amd

    return octree_amd.StaticOrthoTreePointND(D) if kind == "point" else octree_amd.StaticOrthoTreeBoxND(D)


def _export(t):
    from octree_amd._abi import OtgError

    e = t.export(sorted_keys=False)
    try:
        keys = t.export()["sorted_keys"]
    except OtgError:
        keys = None
    return e, keys


def _same_table(a, b):
    for f in ("depth", "ebegin", "ecount", "cbegin", "ccount", "ids", "key"):
        assert np.array_equal(a[f], b[f]), f
    for f in ("min", "size"):
        assert np.array_equal(np.asarray(a[f]).view(np.uint64), np.asarray(b[f]).view(np.uint64)), f


def _same_queries(t0, t1, kind, D, geom, rng):
    nq = 3000
    qs = rng.uniform(geom.min(), geom.max(), (nq, D))
    for k in (1, 8, 33):
        o0, i0 = t0.knn_batch(qs, k).csr()
        o1, i1 = t1.knn_batch(qs, k).csr()
        assert np.array_equal(o0, o1) and np.array_equal(i0, i1), ("knn", k)
    c = qs
    e = rng.uniform(0.5, 8.0, (nq, 1))
    rq = np.concatenate([c - e, c + e], axis=1)
    for mode in (0, 1) if kind == "box" else (0,):
        a = t0.range_batch(rq, mode).csr()
        b = t1.range_batch(rq, mode).csr()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), ("range", mode)
    rays = np.concatenate([qs, unit_dirs(nq, D, seed=5)], axis=1)
    a = t0.ray_first_batch(rays, 0.5 if kind == "point" else 1e-12).csr()
    b = t1.ray_first_batch(rays, 0.5 if kind == "point" else 1e-12).csr()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), "ray"


@pytest.mark.parametrize("kind,D,n,M,cap", [("point", 3, 20_000, 0, 16), ("point", 2, 5_000, 6, 4),
                                            ("point", 6, 8_000, 3, 32), ("box", 3, 12_000, 0, 16),
                                            ("box", 2, 4_000, 5, 8)])
def test_round_trip_table_bytes_and_queries(gpu_ctx, kind, D, n, M, cap):
    print(f"archive load {kind} D={D} n={n} M={M} cap={cap}", flush=True)
    rng = np.random.default_rng(D * 7 + n)
    geom = uniform_points(n, D, 0.0, 100.0, seed=n) if kind == "point" else lognormal_boxes(n, D, 0.0, 100.0, seed=n)
    T = _cls(kind, D)
    t0 = T(ctx=gpu_ctx)
    t0.Create(geom, M or None, None, cap)
    for large in (False, True):
        a = t0.serialize(large_dataset=large)
        t1 = T.deserialize(a, geom, large_dataset=large, ctx=gpu_ctx)
        assert t1.serialize(large_dataset=large) == a
        e0, k0 = _export(t0)
        e1, k1 = _export(t1)
        _same_table(e0, e1)
        if kind == "box":
            assert np.array_equal(k0, k1)
        else:
            assert k1 is None
        _same_queries(t0, t1, kind, D, geom, rng)
        t1.free()
    t0.free()


def test_device_entities_managed_and_forest_sets(gpu_ctx):
    import torch

    import octree_amd

    sets = [gaussian_mixture(m, 3, 4, 0.0, 1000.0, 2.0, 20.0, seed=60 + i) for i, m in enumerate((900, 30_000, 7))]
    f = octree_amd.StaticOctreePoint.forest(sets, 16, None, None, ctx=gpu_ctx)
    for s in range(3):
        a = f.serialize(s)
        # from a device tensor
        t1 = octree_amd.StaticOctreePoint.deserialize(a, torch.from_numpy(sets[s]).cuda(), ctx=gpu_ctx)
        assert t1.serialize() == a
        t2 = octree_amd.StaticOctreePoint(ctx=gpu_ctx)
        t2.Create(sets[s], f.GetMaxDepthID(s), None, 16)
        _same_table(_export(t1)[0], f.export(s, sorted_keys=False))
        # an OrthoTreeManaged archive carries its entities
        m = f.serialize(s, entities=sets[s])
        t3 = octree_amd.StaticOctreePoint.deserialize(m, None, ctx=gpu_ctx)
        assert t3.serialize(entities=sets[s]) == m
        q = sets[s][: min(500, len(sets[s]))] + 0.25
        assert np.array_equal(t3.knn_batch(q, 5).csr()[1], t2.knn_batch(q, 5).csr()[1])
        for t in (t1, t2, t3):
            t.free()
    f.free()


def test_overfull_leaves_get_the_finer_index(gpu_ctx):
    """Clustered points at a shallow depth: max-depth leaves overflow, the loaded tree
    builds the finer kNN / range index and answers as the archived tree."""
    import octree_amd

    pts = gaussian_mixture(60_000, 3, 6, 0.0, 100.0, 0.05, 0.3, seed=9)
    t0 = octree_amd.StaticOctreePoint(ctx=gpu_ctx)
    t0.Create(pts, 3, None, 8)
    a = t0.serialize()
    t1 = octree_amd.StaticOctreePoint.deserialize(a, pts, ctx=gpu_ctx)
    assert t1.info().device_bytes > 60_000 * (3 * 8 + 4) * 1.5  # the index is resident too
    _same_queries(t0, t1, "point", 3, pts, np.random.default_rng(3))
    t0.free()
    t1.free()


def test_reference_round_trip_and_hand_archives(gpu_ctx):
    import octree_amd

    # serialize.tests.cpp:133-146: two points in [0,1]^3, depth 3, then RangeSearch(box)
    pts = np.array([[0.1, 0.1, 0.1], [0.9, 0.9, 0.9]])
    t0 = octree_amd.StaticOctreePoint(ctx=gpu_ctx)
    t0.Create(pts, 3, [0, 0, 0, 1, 1, 1])
    t1 = octree_amd.StaticOctreePoint.deserialize(t0.serialize(entities=pts), None, ctx=gpu_ctx)
    box = [[0.0, 0.0, 0.0, 1.0, 1.0, 1.0]]
    assert sorted(t1.range_batch(box).csr()[1].tolist()) == sorted(t0.range_batch(box).csr()[1].tolist()) == [0, 1]
    # the hand-derived archives (tests/helpers.py) load as the trees they describe
    p = octree_amd.StaticQuadtreePoint.deserialize(hand_archive_point_2d(), np.array(HAND_POINTS_2D), ctx=gpu_ctx)
    e = p.export(sorted_keys=False)
    assert e["depth"].tolist() == [0, 2, 2, 2] and e["ccount"].tolist() == [3, 0, 0, 0]
    assert p.knn_batch([[3.4, 0.6]], 1).csr()[1].tolist() == [1]
    assert p.serialize() == hand_archive_point_2d()
    b = octree_amd.StaticQuadtreeBox.deserialize(hand_archive_box_2d(), np.array(HAND_BOXES_2D), ctx=gpu_ctx)
    assert b.serialize() == hand_archive_box_2d()
    assert sorted(b.range_batch([[0.5, 0.5, 3.5, 3.5]], 1).csr()[1].tolist()) == [0, 1]
    for t in (t0, t1, p, b):
        t.free()


def test_malformed_archives_are_rejected(gpu_ctx):
    import octree_amd
    from octree_amd._abi import OTG_ERR_INVALID_ARG, OtgError

    good = hand_archive_point_2d()
    pts = np.array(HAND_POINTS_2D)
    T = octree_amd.StaticQuadtreePoint

    def bad(buf, ents=pts):
        with pytest.raises(OtgError) as ex:
            T.deserialize(buf, ents, ctx=gpu_ctx)
        assert ex.value.status == OTG_ERR_INVALID_ARG

    bad(good[:-3])                                  # truncated
    bad(good + b"\0")                               # trailing bytes
    bad(struct.pack("<I", 1) + good[4:])            # unknown version
    ids_at = good.index(struct.pack("<I", 3) + struct.pack("<3I", 0, 1, 2))
    dup = good[:ids_at + 4] + struct.pack("<3I", 0, 0, 2) + good[ids_at + 16:]
    bad(dup)                                        # ids not a permutation
    seg_at = good.index(bytes([0, 0, 0, 1, 1, 1, 2, 1]))
    overlap = good[:seg_at] + bytes([0, 0, 0, 1, 0, 1, 2, 1]) + good[seg_at + 8:]
    bad(overlap)                                    # children's entities not contiguous
    bad(good, pts[[1, 0, 2]])                       # other entities: a point outside its leaf
    bad(good, pts[:2])                              # entity count differs
    assert T.deserialize(good, pts, ctx=gpu_ctx).serialize() == good
