"""Create<SeqExec> on the GPU (otg_tree_build_seq): the node numbering and entity
order of the reference's unsorted build (Build<false>, ot_static_linear_core.h:417-519)
must equal the oracle's bit for bit -- node records, MinPoint / MBR geometry, the
entity storage order and the keys in that order -- and queries on the SEQ tree must
answer as on the sorted (PAR) tree."""
import numpy as np
import pytest

from helpers import gaussian_mixture, lognormal_boxes, uniform_points, unit_dirs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _seq_tree(ctx, kind, D, geom, box, M, cap):
    import octree_amd

    cls = octree_amd.StaticOrthoTreePointND(D) if kind == "point" else octree_amd.StaticOrthoTreeBoxND(D)
    t = cls(ctx=ctx)
    t.Create(geom, M or None, box, cap, exec_tag=octree_amd.SEQ_EXEC)
    return t


def _same(g, o, kind):
    for f in ("depth", "ebegin", "ecount", "ccount", "key", "ids"):
        assert np.array_equal(g[f], o[f]), f
    m = o["ccount"] > 0
    assert np.array_equal(g["cbegin"][m], o["cbegin"][m]), "cbegin"
    for f in ("min", "size"):
        bad = np.nonzero((g[f].view(np.uint64) != o[f].view(np.uint64)).any(axis=1))[0]
        assert len(bad) == 0, (f, bad[:5])
    if kind == "box":
        db = max(1, int(o["max_depth"]).bit_length())
        ok = (o["sorted_code"].astype(np.uint64) << np.uint64(db)) | o["sorted_depth"].astype(np.uint64)
    else:
        ok = o["sorted_code"]
    assert np.array_equal(g["sorted_keys"], ok), "keys in storage order"


@pytest.mark.parametrize("kind,D,gen,n,M,cap", [
    ("point", 3, "uniform", 20_000, 0, 16), ("point", 2, "uniform", 5_000, 7, 1), ("point", 6, "cluster", 20_000, 3, 16),
    ("point", 3, "dups", 8_000, 6, 4), ("point", 4, "cluster", 30_000, 0, 8), ("box", 3, "boxes", 20_000, 0, 16),
    ("box", 2, "boxes", 6_000, 6, 2), ("box", 3, "boxes", 3_000, 4, 64)])
def test_seq_table_vs_oracle(gpu_ctx, kind, D, gen, n, M, cap):
    print(f"seq build {kind} D={D} {gen} n={n} M={M} cap={cap}", flush=True)
    rng = np.random.default_rng(D * 100 + n)
    if gen == "uniform":
        geom = uniform_points(n, D, seed=n)
    elif gen == "cluster":
        geom = gaussian_mixture(n, D, 6, 0.0, 1.0, 0.005, 0.05, seed=D)
    elif gen == "dups":
        geom = uniform_points(n, D, seed=5)
        geom[: n // 8] = geom[rng.integers(n // 8, n, n // 8)]
    else:
        geom = lognormal_boxes(n, D, 0.0, 100.0, seed=n)
    t = _seq_tree(gpu_ctx, kind, D, geom, None, M, cap)
    o = O.OracleTree(D, kind, geom, None, M, cap, seq=True)
    _same(t.export(), o.export(), kind)
    t.free()


def test_seq_c1_size_and_queries(gpu_ctx):
    """The 1M-point octree of the C1 workload (depth 10, cap 16) built SEQ equals the
    oracle's SEQ tree, and its kNN / range / ray answers equal the PAR tree's."""
    import octree_amd

    pts = uniform_points(1_000_000, 3, 0.0, 100.0, seed=2024)
    box = [0, 0, 0, 100, 100, 100]
    s = _seq_tree(gpu_ctx, "point", 3, pts, box, 10, 16)
    _same(s.export(), O.OracleTree(3, "point", pts, box, 10, 16, seq=True).export(), "point")
    p = octree_amd.StaticOctreePoint(ctx=gpu_ctx)
    p.Create(pts, 10, box, 16)
    qs = uniform_points(20_000, 3, 0.0, 100.0, seed=2025)
    for k in (1, 8):
        a, b = s.knn_batch(qs, k).csr(), p.knn_batch(qs, k).csr()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    rq = np.concatenate([qs - 1.5, qs + 1.5], axis=1)
    a, b = s.range_batch(rq).csr(), p.range_batch(rq).csr()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    rays = np.concatenate([qs, unit_dirs(len(qs), 3, seed=3)], axis=1)
    a, b = s.ray_first_batch(rays, 0.05).csr(), p.ray_first_batch(rays, 0.05).csr()
    assert np.array_equal(a[0], b[0])
    s.free()
    p.free()


def test_seq_box_queries_and_archive(gpu_ctx):
    import octree_amd

    boxes = lognormal_boxes(40_000, 3, 0.0, 1000.0, seed=4)
    s = _seq_tree(gpu_ctx, "box", 3, boxes, None, 0, 16)
    p = octree_amd.StaticOctreeBox(ctx=gpu_ctx)
    p.Create(boxes, None, None, 16)
    assert sorted(map(tuple, s.collide())) == sorted(map(tuple, p.collide()))
    q = np.random.default_rng(1).uniform(0, 1000, (5000, 3))
    rq = np.concatenate([q - 10, q + 10], axis=1)
    for mode in (0, 1):
        a, b = s.range_batch(rq, mode).csr(), p.range_batch(rq, mode).csr()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # the SEQ tree archives and loads back like any other
    arc = s.serialize()
    t = octree_amd.StaticOctreeBox.deserialize(arc, boxes, ctx=gpu_ctx)
    assert t.serialize() == arc
    for x in (s, p, t):
        x.free()


def test_seq_empty_and_single(gpu_ctx):
    for n in (0, 1, 2):
        geom = uniform_points(n, 3, seed=1)
        t = _seq_tree(gpu_ctx, "point", 3, geom, [0, 0, 0, 100, 100, 100], 4, 1)
        o = O.OracleTree(3, "point", geom, [0, 0, 0, 100, 100, 100], 4, 1, seq=True)
        _same(t.export(), o.export(), "point")
        t.free()
