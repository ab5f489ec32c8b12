"""GPU parity of the batched queries against the oracle and the reference's KATs:
kNN (ot_query.h:1208-1270), RangeSearch (:476-583), RayIntersectedFirst
(:2035-2122), CollisionDetection (:1901-1906).  Ids must be identical; kNN and
ray lists are compared in order (distance ties ordered by id on both sides)."""
import numpy as np
import pytest

from helpers import (BASE_TOL, DBL_MAX, MT19937, collision_brute, gaussian_mixture, issue38_boxes, kat, knn_brute,
                     lognormal_boxes, uniform_points, unit_dirs)
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _tree(ctx, kind, D, geom, box, M, cap):
    import octree_amd

    cls = octree_amd.StaticOrthoTreePointND(D) if kind == "point" else octree_amd.StaticOrthoTreeBoxND(D)
    t = cls(ctx=ctx)
    t.Create(geom, M or None, box, cap)
    return t


# ---- kNN ----------------------------------------------------------------------------
@pytest.mark.parametrize("D,n,k,cap,M", [(2, 3000, 4, 5, 10), (3, 20000, 8, 16, 7), (3, 5000, 1, 2, 8),
                                         (6, 20000, 16, 32, 5), (3, 4000, 32, 8, 6), (3, 500, 64, 4, 5)])
def test_knn_vs_oracle(gpu_ctx, D, n, k, cap, M):
    pts = uniform_points(n, D, seed=D + n + k)
    pts[:100] = pts[100:200]
    qs = uniform_points(2000, D, seed=7 + D)
    qs[:50] = pts[:50]
    g = _tree(gpu_ctx, "point", D, pts, None, M, cap)
    o = O.OracleTree(D, "point", pts, None, M, cap)
    goff, gids = g.knn_batch(qs, k).csr()
    ooff, oids = o.knn(qs, k, threads=8)
    assert np.array_equal(goff, ooff)
    assert np.array_equal(gids, oids)


@pytest.mark.parametrize("D,k", [(4, 8), (5, 16), (6, 16), (6, 1)])
def test_knn_wide_dims_clustered(gpu_ctx, D, k):
    """4-D..6-D walks (one child mask per level): clustered points under a shallow
    depth, so leaves overflow and the finer kNN index is walked, equal the oracle."""
    pts = gaussian_mixture(20000, D, 8, 0.0, 1.0, 0.01, 0.05, seed=80 + D)
    rng = np.random.default_rng(90 + D)
    qs = pts[rng.integers(0, len(pts), 2000)] + rng.normal(0.0, 0.01, (2000, D))
    g = _tree(gpu_ctx, "point", D, pts, None, 3, 8)
    o = O.OracleTree(D, "point", pts, None, 3, 8)
    goff, gids = g.knn_batch(qs, k).csr()
    ooff, oids = o.knn(qs, k, threads=8)
    assert np.array_equal(goff, ooff)
    assert np.array_equal(gids, oids)


def test_knn_max_within_and_tolerance(gpu_ctx):
    pts = uniform_points(5000, 3, seed=1)
    qs = uniform_points(1000, 3, seed=2)
    g = _tree(gpu_ctx, "point", 3, pts, None, 6, 8)
    o = O.OracleTree(3, "point", pts, None, 6, 8)
    for mw, tol in ((0.0, BASE_TOL), (1.5, BASE_TOL), (4.0, 0.0), (DBL_MAX, 1e-6)):
        goff, gids = g.knn_batch(qs, 8, mw, tol).csr()
        ooff, oids = o.knn(qs, 8, mw, tol, threads=8)
        assert np.array_equal(goff, ooff) and np.array_equal(gids, oids), (mw, tol)


def test_knn_kats(gpu_ctx):
    k = kat()["knn_point_set_no1"]
    g = _tree(gpu_ctx, "point", 2, np.array(k["points"]), None, k["max_depth"], k["cap"])
    for c in k["cases"]:
        got = g.GetNearestNeighbors(c["q"], c["k"], c.get("max_within", DBL_MAX), None, c.get("tol", BASE_TOL))
        assert sorted(got) == sorted(c["expect"]), c["src"]
    for c in kat()["knn_issue9"]:
        pts = np.array(c["points"])
        g = _tree(gpu_ctx, "point", c["dim"], pts, c["box"], c["max_depth"], c["cap"])
        d2 = ((pts - np.array(c["q"])) ** 2).sum(axis=1)
        assert g.GetNearestNeighbors(c["q"], 1)[0] == int(np.argmin(d2)), c["src"]


def test_knn_c1_size_brute_sample(gpu_ctx):
    """1M points, 10k queries k=8: GPU == oracle on all, == brute force on a sample."""
    pts = uniform_points(1_000_000, 3, 0.0, 100.0, seed=2024)
    qs = uniform_points(10_000, 3, 0.0, 100.0, seed=2025)
    box = [0, 0, 0, 100, 100, 100]
    g = _tree(gpu_ctx, "point", 3, pts, box, 10, 16)
    o = O.OracleTree(3, "point", pts, box, 10, 16)
    goff, gids = g.knn_batch(qs, 8).csr()
    ooff, oids = o.knn(qs, 8, threads=16)
    assert np.array_equal(goff, ooff) and np.array_equal(gids, oids)
    for qi in range(0, 10_000, 500):
        exp, amb = knn_brute(qs[qi], pts, 8)
        assert not amb
        assert gids[goff[qi]:goff[qi + 1]].tolist() == exp


def test_knn_band_longer_than_list(gpu_ctx):
    """Many entities inside the tolerance band (duplicates at the query): the list
    pass overflows and the band pass returns every id below F, sorted."""
    import octree_amd

    pts = uniform_points(6000, 3, seed=31)
    pts[:300] = pts[5000]  # 300 + 1 identical points
    qs = np.concatenate([pts[5000:5001], uniform_points(500, 3, seed=32), pts[:3] + 1e-15])
    g = _tree(gpu_ctx, "point", 3, pts, None, 6, 8)
    o = O.OracleTree(3, "point", pts, None, 6, 8)
    ctx = gpu_ctx
    ctx.set_profiling(True)
    ctx.reset_kernel_times()
    for k in (1, 5, 40):
        goff, gids = g.knn_batch(qs, k).csr()
        ooff, oids = o.knn(qs, k, threads=8)
        assert np.array_equal(goff, ooff) and np.array_equal(gids, oids), k
        assert goff[1] - goff[0] == 301
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    assert "knn_band_fill" in kt, sorted(kt)


def test_knn_6d_limits_ties_and_band(gpu_ctx):
    """6-D point kNN (one wavefront per query): distance limits and tolerances, equal
    distances ordered by id, and tolerance bands longer than the list, against the oracle."""
    pts = uniform_points(20000, 6, seed=61)
    pts[100:103] = pts[7000]        # three copies: equal distances inside a list
    pts[200:400] = pts[9000]        # 200 + 1 copies: the band outgrows any list
    qs = np.concatenate([pts[7000:7001] + 1e-3, pts[9000:9001], uniform_points(1500, 6, seed=62)])
    g = _tree(gpu_ctx, "point", 6, pts, None, 4, 16)
    o = O.OracleTree(6, "point", pts, None, 4, 16)
    for k, mw, tol in ((8, DBL_MAX, BASE_TOL), (1, DBL_MAX, BASE_TOL), (20, DBL_MAX, BASE_TOL), (8, 0.5, BASE_TOL),
                       (16, 12.0, 0.0), (5, DBL_MAX, 1e-6)):
        goff, gids = g.knn_batch(qs, k, mw, tol).csr()
        ooff, oids = o.knn(qs, k, mw, tol, threads=8)
        assert np.array_equal(goff, ooff) and np.array_equal(gids, oids), (k, mw, tol)
    assert goff[2] - goff[1] == 201  # the band pass returned every copy


# ---- box-tree kNN (MinMaxDistance) -----------------------------------------------------
def test_box_knn_kats(gpu_ctx):
    for c in kat()["knn_box_2d"]:
        g = _tree(gpu_ctx, "box", 2, np.array(c["boxes"]), None, c["depth"], c["cap"])
        off, ids = g.knn_batch([c["q"]], c["k"], DBL_MAX, c.get("tol", BASE_TOL)).csr()
        assert sorted(ids.tolist()) == sorted(c["expect"]), c["src"]


@pytest.mark.parametrize("D,n,k,cap,M", [(2, 5000, 1, 4, 6), (3, 20000, 8, 16, 6), (3, 8000, 16, 8, 5),
                                         (3, 3000, 40, 8, 5), (4, 6000, 4, 8, 4), (6, 4000, 4, 16, 3)])
def test_box_knn_vs_oracle(gpu_ctx, D, n, k, cap, M):
    boxes = lognormal_boxes(n, D, 0.0, 100.0, np.log(0.8), 0.7, seed=D * 3 + k)
    qs = uniform_points(1500, D, seed=5 + D + k)
    qs[:10] = boxes[:10, :D]  # on box corners
    g = _tree(gpu_ctx, "box", D, boxes, None, M, cap)
    o = O.OracleTree(D, "box", boxes, None, M, cap)
    goff, gids = g.knn_batch(qs, k).csr()
    ooff, oids = o.knn(qs, k, threads=8)
    assert np.array_equal(goff, ooff)
    assert np.array_equal(gids, oids)


# ---- range ----------------------------------------------------------------------------
@pytest.mark.parametrize("kind,D,mode", [("point", 2, 0), ("point", 3, 0), ("point", 6, 0), ("box", 3, 0), ("box", 3, 1),
                                         ("box", 2, 1)])
def test_range_vs_oracle(gpu_ctx, kind, D, mode):
    n = 20000
    if kind == "point":
        geom = uniform_points(n, D, 0.0, 100.0, seed=D)
    else:
        geom = lognormal_boxes(n, D, 0.0, 100.0, np.log(0.5), 0.8, seed=D)
    rng = np.random.default_rng(D + mode)
    c = rng.uniform(0, 100, (3000, D))
    e = rng.uniform(0.5, 15, (3000, 1))
    ranges = np.concatenate([c - e, c + e], axis=1)
    ranges[0] = np.concatenate([np.full(D, -5.0), np.full(D, 105.0)])
    ranges[1] = np.concatenate([np.full(D, 5.0), np.full(D, 5.0)])
    ranges[2] = np.concatenate([np.full(D, 0.0), np.full(D, 50.0)])
    g = _tree(gpu_ctx, kind, D, geom, None, 6, 8)
    o = O.OracleTree(D, kind, geom, None, 6, 8)
    goff, gids = g.range_batch(ranges, mode).csr()
    ooff, oids = o.range(ranges, overlap=bool(mode), threads=8)
    assert np.array_equal(goff, ooff)
    for qi in range(len(ranges)):
        # each list sorted by id; the oracle's is the reference's walk order
        assert gids[goff[qi]:goff[qi + 1]].tolist() == sorted(oids[ooff[qi]:ooff[qi + 1]].tolist()), qi


@pytest.mark.parametrize("records", [0, 1, 3, 64])
def test_range_kept_records(gpu_ctx, records):
    """Large ranges whose lists the wavefront count walks keep as records: with the
    default slots, with 1 or 3 slots per query (most records in the shared overflow
    area, which then runs out: those queries take the fill walk) and with 64."""
    n, D = 200_000, 3
    geom = gaussian_mixture(n, D, 5, 0.0, 100.0, 0.5, 6.0, seed=41)
    rng = np.random.default_rng(42)
    c = rng.uniform(0, 100, (1500, D))
    e = rng.uniform(2.0, 30.0, (1500, 1))
    ranges = np.concatenate([c - e, c + e], axis=1)
    ranges[0] = np.concatenate([np.full(D, -5.0), np.full(D, 105.0)])  # the whole set
    o = O.OracleTree(D, "point", geom, None, 8, 8)
    ooff, oids = o.range(ranges, overlap=False, threads=8)
    g = _tree(gpu_ctx, "point", D, geom, None, 8, 8)
    try:
        gpu_ctx.set_tuning(range_records=records)
        goff, gids = g.range_batch(ranges, 0).csr()
    finally:
        gpu_ctx.set_tuning()
    assert np.array_equal(goff, ooff)
    for qi in range(len(ranges)):
        assert gids[goff[qi]:goff[qi + 1]].tolist() == sorted(oids[ooff[qi]:ooff[qi + 1]].tolist()), qi


# ---- ray first -----------------------------------------------------------------------
def test_ray_kats(gpu_ctx):
    for c in kat()["ray_first_2d"]:
        g = _tree(gpu_ctx, "box", 2, np.array(c["boxes"]), None, c["depth"], c["cap"])
        got = g.RayIntersectedFirst(c["o"], c["d"], None, c["tol"], c.get("inc", 0.0), c.get("max_dist", DBL_MAX))
        assert got == c["expect"], c["src"]


@pytest.mark.parametrize("kind,D", [("box", 3), ("box", 2), ("point", 3)])
def test_ray_vs_oracle(gpu_ctx, kind, D):
    n = 30000
    if kind == "box":
        geom = lognormal_boxes(n, D, 0.0, 1000.0, np.log(0.5 * 4), 0.6, seed=21)
    else:
        geom = uniform_points(n, D, 0.0, 1000.0, seed=21)
    o_ = uniform_points(3000, D, 0.0, 1000.0, seed=22)
    d_ = unit_dirs(3000, D, seed=23)
    rays = np.concatenate([o_, d_], axis=1)
    rays[0, D:] = 0.0
    rays[0, D] = 1.0  # axis-aligned: zero components take the NaN path
    rays[1, D:] = 2.0  # not normalised -> empty (igm:541-545)
    tol = BASE_TOL if kind == "box" else 0.5
    g = _tree(gpu_ctx, kind, D, geom, None, 7, 8)
    o = O.OracleTree(D, kind, geom, None, 7, 8)
    goff, gids = g.ray_first_batch(rays, tol).csr()
    ooff, oids = o.ray_first(rays, tol, threads=8)
    assert np.array_equal(goff, ooff)
    assert np.array_equal(gids, oids)
    assert goff[2] - goff[1] == 0


def test_ray_points_clustered_wide_tolerance(gpu_ctx):
    """Clustered points under a shallow max depth (over-full leaves) with wide
    tolerance cones: each hit can raise the bound M, so the result depends on the
    order hits are met; rays walk the reference tree nearest cell first, which is
    the reference's best-first order for non-overlapping cells (2-D and 3-D)."""
    for D in (2, 3):
        pts = gaussian_mixture(40000, D, 5, 0.0, 100.0, 0.5, 3.0, seed=50 + D)
        g = _tree(gpu_ctx, "point", D, pts, None, 3, 8)
        o = O.OracleTree(D, "point", pts, None, 3, 8)
        org = pts[np.random.default_rng(D).integers(0, len(pts), 2000)] + 2.0
        rays = np.concatenate([org, unit_dirs(2000, D, seed=60 + D)], axis=1)
        for tol in (0.05, 0.3):
            goff, gids = g.ray_first_batch(rays, tol).csr()
            ooff, oids = o.ray_first(rays, tol, threads=8)
            assert np.array_equal(goff, ooff) and np.array_equal(gids, oids), (D, tol)
            assert len(gids) > 100


def test_ray_band_longer_than_list(gpu_ctx):
    """Points strung along the ray inside its tolerance cone: more first-hit
    candidates than the per-ray list; the exact replay with a larger list returns them."""
    line = np.zeros((200, 3))
    line[:, 0] = 10.0 + 0.01 * np.arange(200)
    pts = np.concatenate([uniform_points(5000, 3, seed=71), line])
    g = _tree(gpu_ctx, "point", 3, pts, None, 6, 8)
    o = O.OracleTree(3, "point", pts, None, 6, 8)
    rays = np.array([[0.0, 0.0, 0.0, 1.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0, 1.0, 0.0]])
    gpu_ctx.set_profiling(True)
    gpu_ctx.reset_kernel_times()
    goff, gids = g.ray_first_batch(rays, 0.5).csr()
    kt = gpu_ctx.kernel_times()
    gpu_ctx.set_profiling(False)
    ooff, oids = o.ray_first(rays, 0.5, threads=8)
    assert np.array_equal(goff, ooff) and np.array_equal(gids, oids)
    assert goff[1] - goff[0] > 16 and "ray_replay" in kt, (goff, sorted(kt))


def test_ray_frontier_beyond_private_heaps(gpu_ctx):
    """Rays from far outside a dense tree with tolerance cones that widen with
    distance (increment 0.2) reach every node, best-first in enter order: the queued
    frontier becomes a slab of the tree, wider than the LDS heap and the private
    heap, and every entity is a candidate.  The walk then runs with a node heap of
    the tree's size in global memory and the replay grows its candidate list to the
    set's size; the ids equal the oracle's.  A tolerance that hits everything at
    enter distance 0 (a depth-first frontier) is checked too."""
    D, n = 3, 30000
    pts = uniform_points(n, D, seed=83)
    g = _tree(gpu_ctx, "point", D, pts, None, 8, 2)
    o = O.OracleTree(D, "point", pts, None, 8, 2)
    org = np.array([[-1000.0, 50.0, 50.0], [50.0, -1000.0, 40.0], [-600.0, -700.0, 30.0]])
    dirs = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]] + [list(-org[2] / np.linalg.norm(org[2]))])
    rays = np.concatenate([org, dirs], axis=1)
    for tol, inc, wide in ((1.0, 0.2, True), (1e6, 0.0, False)):
        gpu_ctx.set_profiling(True)
        gpu_ctx.reset_kernel_times()
        goff, gids = g.ray_first_batch(rays, tol, inc).csr()
        kt = gpu_ctx.kernel_times()
        gpu_ctx.set_profiling(False)
        print(f"tolerance {tol}: {int(goff[-1])} ids, kernels {sorted(kt)}", flush=True)
        ooff, oids = o.ray_first(rays, tol, inc, threads=8)
        assert np.array_equal(goff, ooff), (tol, goff, ooff)
        for i in range(len(rays)):  # equal enter distances: compared as sets within each run
            a, b = gids[goff[i]:goff[i + 1]], oids[ooff[i]:ooff[i + 1]]
            assert sorted(a.tolist()) == sorted(b.tolist()), (tol, i)
        assert goff[1] - goff[0] > 500
        if wide:
            assert "ray_deep_global" in kt and "ray_replay_global" in kt, sorted(kt)


# ---- kNN with k above the register lists ---------------------------------------------------
@pytest.mark.parametrize("kind,D,k", [("point", 3, 100), ("point", 2, 257), ("point", 6, 80), ("box", 3, 90)])
def test_knn_k_above_64(gpu_ctx, kind, D, k):
    n = 20000
    if kind == "point":
        geom = gaussian_mixture(n, D, 6, 0.0, 100.0, 1.0, 6.0, seed=D + k)
        geom[:300] = geom[300:600]  # exact duplicates: equal distances
    else:
        geom = lognormal_boxes(n, D, 0.0, 100.0, np.log(0.5), 0.8, seed=D)
    qs = uniform_points(3000, D, 0.0, 100.0, seed=k)
    qs[:30] = (geom[:30, :D] if kind == "point" else geom[:30, :D] + 0.1)
    g = _tree(gpu_ctx, kind, D, geom, None, 6, 16)
    o = O.OracleTree(D, kind, geom, None, 6, 16)
    for mw in (DBL_MAX, 3.0):
        goff, gids = g.knn_batch(qs, k, mw).csr()
        ooff, oids = o.knn(qs, k, mw, threads=8)
        assert np.array_equal(goff, ooff) and np.array_equal(gids, oids), (kind, D, k, mw)
        if mw == DBL_MAX:
            assert (np.diff(goff) >= k).all()


# ---- collision -----------------------------------------------------------------------
def test_collision_issue38(gpu_ctx):
    rng = MT19937(0)
    for _ in range(10):
        b = issue38_boxes(rng)
        g = _tree(gpu_ctx, "box", 3, b, None, 10, 5)
        assert np.array_equal(g.collide().astype(np.int64), collision_brute(b))


def test_collision_vs_oracle(gpu_ctx):
    b = lognormal_boxes(200_000, 3, 0.0, 1000.0, np.log(0.5), 0.6, seed=31)
    b = np.concatenate([b, lognormal_boxes(2000, 3, 0.0, 1000.0, np.log(20.0), 0.5, seed=32)])
    g = _tree(gpu_ctx, "box", 3, b, None, 8, 32)
    o = O.OracleTree(3, "box", b, None, 8, 32)
    assert np.array_equal(g.collide(), o.collide())


# ---- forest queries -------------------------------------------------------------------
def test_forest_queries_match_per_set_trees(gpu_ctx):
    import octree_amd

    D = 3
    sizes = [5, 700, 3000, 40, 9000, 1]
    sets = [gaussian_mixture(n, D, 8, 0.0, 1000.0, 2.0, 20.0, seed=40 + i) for i, n in enumerate(sizes)]
    box = np.tile([[0.0] * D + [1000.0] * D], (len(sizes), 1))
    f = octree_amd.StaticOctreePoint.forest(sets, 8, None, box, ctx=gpu_ctx)
    rng = np.random.default_rng(9)
    nq = 4000
    sid = rng.integers(0, len(sizes), nq).astype(np.uint32)
    qs = uniform_points(nq, D, 0.0, 1000.0, seed=10)
    c = rng.uniform(0, 1000, (nq, D))
    e = rng.uniform(1, 30, (nq, 1))
    rq = np.concatenate([c - e, c + e], axis=1)
    koff, kids = f.knn_batch(qs, 16, set_ids=sid).csr()
    roff, rids = f.range_batch(rq, set_ids=sid).csr()
    for s in range(len(sizes)):
        o = O.OracleTree(D, "point", sets[s], box[s], 0, 8)
        m = np.nonzero(sid == s)[0]
        ooff, oids = o.knn(qs[m], 16)
        for j, qi in enumerate(m):
            assert kids[koff[qi]:koff[qi + 1]].tolist() == oids[ooff[j]:ooff[j + 1]].tolist()
        ooff, oids = o.range(rq[m])
        for j, qi in enumerate(m):
            assert sorted(rids[roff[qi]:roff[qi + 1]].tolist()) == sorted(oids[ooff[j]:ooff[j + 1]].tolist())


# ---- stack-heavy walks ------------------------------------------------------------------
def test_short_stack_paths(gpu_ctx):
    """Walks that keep many siblings pending: far-away 3-D kNN with a large k and
    ranges cutting nodes at every level (one LDS entry per level), and 6-D kNN
    with up to 63 pending siblings per level (private worst-case stack)."""
    import octree_amd

    ctx = octree_amd.Context(0)
    pts = uniform_points(20000, 3, seed=91)
    g = _tree(ctx, "point", 3, pts, None, 9, 1)
    o = O.OracleTree(3, "point", pts, None, 9, 1)
    # far-away kNN with a large k keeps many siblings within the radius
    qs = np.concatenate([uniform_points(200, 3, seed=92) + np.array([400.0, 0.0, 0.0]), uniform_points(300, 3, seed=93)])
    # ranges covering most of the space cut through nodes at every level
    rg = np.array([[1.0, 1.0, 1.0, 99.0, 99.0, 99.0], [0.5, 2.0, 3.0, 97.0, 98.0, 99.5],
                   [10.0, 10.0, 10.0, 60.0, 60.0, 60.0]])
    got_k = g.knn_batch(qs, 64).csr()
    got_r = g.range_batch(rg).csr()
    exp_k = o.knn(qs, 64, threads=8)
    exp_r = o.range(rg, threads=8)
    assert np.array_equal(got_k[0], exp_k[0]) and np.array_equal(got_k[1], exp_k[1])
    assert np.array_equal(got_r[0], exp_r[0])
    for qi in range(len(rg)):  # range output order is the walk's: compared as sets
        a, b = got_r[0][qi], got_r[0][qi + 1]
        assert sorted(got_r[1][a:b].tolist()) == sorted(exp_r[1][a:b].tolist()), qi
    g.free()
    # 6-D: up to 63 pending siblings per level overflow the short stack
    p6 = uniform_points(30000, 6, seed=94)
    g6 = _tree(ctx, "point", 6, p6, None, 4, 2)
    o6 = O.OracleTree(6, "point", p6, None, 4, 2)
    q6 = np.concatenate([uniform_points(100, 6, seed=95) + 300.0, uniform_points(200, 6, seed=96)])
    got6 = g6.knn_batch(q6, 32).csr()
    exp6 = o6.knn(q6, 32, threads=8)
    assert np.array_equal(got6[0], exp6[0]) and np.array_equal(got6[1], exp6[1])
    g6.free()
    ctx.close()
