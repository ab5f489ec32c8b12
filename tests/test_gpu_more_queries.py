"""GPU parity of PickSearch (ot_query.h:426-466), RayIntersectedAll (:1929-2022)
and two-tree CollisionDetection (:1462-1496) against the oracle and the
reference's KATs.  Pick lists are compared as sets (the reference's output
order is its walk order, and ids of equal location keys have no defined order);
RayIntersectedAll lists in order (ascending enter distance, ties by id on both
sides); collision pairs (left, right) sorted."""
import numpy as np
import pytest

from helpers import BASE_TOL, DBL_MAX, collision_two_brute, gaussian_mixture, kat, lognormal_boxes, uniform_points, \
    unit_dirs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _tree(ctx, kind, D, geom, box, M, cap):
    import octree_amd

    cls = octree_amd.StaticOrthoTreePointND(D) if kind == "point" else octree_amd.StaticOrthoTreeBoxND(D)
    t = cls(ctx=ctx)
    t.Create(geom, M or None, box, cap)
    return t


def _same_sets(goff, gids, ooff, oids):
    assert np.array_equal(goff, ooff)
    for qi in range(len(goff) - 1):
        assert sorted(gids[goff[qi]:goff[qi + 1]].tolist()) == sorted(oids[ooff[qi]:ooff[qi + 1]].tolist()), qi


# ---- PickSearch ----------------------------------------------------------------------------
def test_pick_kats(gpu_ctx):
    for c in kat()["pick"]:
        if "boxes" in c:
            g = _tree(gpu_ctx, "box", c["dim"], np.array(c["boxes"], float), None, c["depth"], c["cap"])
        else:
            g = _tree(gpu_ctx, "point", c["dim"], np.array(c["points"], float), None, c["depth"], c["cap"])
        got = g.PickSearch(c["p"], None, c.get("tol", BASE_TOL))
        assert sorted(got) == sorted(c["expect"]), c["src"]


@pytest.mark.parametrize("kind,D,tol", [("box", 3, BASE_TOL), ("box", 2, 0.0), ("box", 3, 0.5), ("box", 6, BASE_TOL),
                                        ("point", 3, 0.06), ("point", 2, 0.0)])
def test_pick_vs_oracle(gpu_ctx, kind, D, tol):
    n = 20000
    if kind == "box":
        geom = lognormal_boxes(n, D, 0.0, 100.0, np.log(3.0), 0.6, seed=60 + D)
        picks = uniform_points(3000, D, 0.0, 100.0, seed=61)
        picks[:500] = geom[:500, :D]         # on min corners
        picks[500:800] = geom[500:800, D:]   # on max corners
    else:
        geom = np.round(uniform_points(n, D, 0.0, 100.0, seed=62 + D), 1)  # exact duplicates
        picks = np.concatenate([geom[:1500], uniform_points(1500, D, 0.0, 100.0, seed=63)])
    g = _tree(gpu_ctx, kind, D, geom, None, 6, 8)
    o = O.OracleTree(D, kind, geom, None, 6, 8)
    goff, gids = g.pick_batch(picks, tol).csr()
    ooff, oids = o.pick(picks, tol, threads=8)
    _same_sets(goff, gids, ooff, oids)
    assert goff[-1] > 0


# ---- RayIntersectedAll -----------------------------------------------------------------------
def test_ray_all_kats(gpu_ctx):
    for c in kat()["ray_all"]:
        g = _tree(gpu_ctx, "box", c["dim"], np.array(c["boxes"], float), None, c["depth"], c["cap"])
        got = g.RayIntersectedAll(c["o"], c["d"], None, c["tol"], c.get("inc", 0.0), c.get("max_dist", DBL_MAX))
        if c.get("ordered"):
            assert got == c["expect"], c["src"]
        else:
            assert sorted(got) == sorted(c["expect"]), c["src"]


@pytest.mark.parametrize("kind,D,tol,inc,max_dist", [("box", 3, BASE_TOL, 0.0, DBL_MAX), ("box", 2, 0.0, 0.0, 300.0),
                                                     ("box", 3, 1e-3, 1e-3, 0.0), ("box", 5, BASE_TOL, 0.0, DBL_MAX),
                                                     ("point", 3, 0.5, 0.0, DBL_MAX), ("point", 3, 0.2, 1e-3, 500.0)])
def test_ray_all_vs_oracle(gpu_ctx, kind, D, tol, inc, max_dist):
    n = 30000
    hi = 1000.0 if D <= 3 else 100.0  # boxes of ~6 per side in [0, 100]^5: ~10 hits per ray
    if kind == "box":
        geom = lognormal_boxes(n, D, 0.0, hi, np.log(2.0 if D <= 3 else 6.0), 0.6, seed=70 + D)
    else:
        geom = gaussian_mixture(n, D, 8, 0.0, hi, 5.0, 30.0, seed=71)
    rays = np.concatenate([uniform_points(3000, D, 0.0, hi, seed=72), unit_dirs(3000, D, seed=73)], axis=1)
    rays[0, D:] = 0.0
    rays[0, D] = 1.0   # axis-aligned: zero components take the NaN path
    rays[1, D:] = 2.0  # not normalised -> empty
    g = _tree(gpu_ctx, kind, D, geom, None, 7, 8)
    o = O.OracleTree(D, kind, geom, None, 7, 8)
    goff, gids = g.ray_all_batch(rays, tol, inc, max_dist).csr()
    ooff, oids = o.ray_all(rays, tol, inc, max_dist, threads=8)
    assert np.array_equal(goff, ooff)
    assert np.array_equal(gids, oids)
    assert goff[2] - goff[1] == 0 and goff[-1] > 0


# ---- two-tree CollisionDetection ------------------------------------------------------------------
def test_collision_two_trees_kats(gpu_ctx):
    import octree_amd

    for c in kat()["collision_two_trees"]:
        gl = _tree(gpu_ctx, "box", 2, np.array(c["left"], float), None, c["depth"], c["cap"])
        gr = _tree(gpu_ctx, "box", 2, np.array(c["right"], float), None, c["depth"], c["cap"])
        got = sorted(octree_amd.StaticTree.CollisionDetectionTrees(gl, None, gr, None))
        assert got == sorted(map(tuple, c["expect"])), c["src"]
        assert sorted(gl.CollisionDetection(None, gr, None)) == got


@pytest.mark.parametrize("D", [2, 3])
def test_collision_two_trees_vs_oracle(gpu_ctx, D):
    left = lognormal_boxes(20000, D, 0.0, 200.0, np.log(0.8), 0.6, seed=80 + D)
    right = lognormal_boxes(12000, D, 50.0, 250.0, np.log(0.8), 0.6, seed=81 + D)
    gl, gr = _tree(gpu_ctx, "box", D, left, None, 6, 8), _tree(gpu_ctx, "box", D, right, None, 7, 4)
    ol, orr = O.OracleTree(D, "box", left, None, 6, 8), O.OracleTree(D, "box", right, None, 7, 4)
    got = gl.collide_with(gr).astype(np.int64)
    exp = ol.collide_with(orr).astype(np.int64)
    assert len(exp) > 0
    assert np.array_equal(got, exp)
    back = gr.collide_with(gl).astype(np.int64)
    assert np.array_equal(back[np.lexsort((back[:, 0], back[:, 1]))][:, ::-1], got)
    if D == 2:
        assert np.array_equal(got, collision_two_brute(left, right))


def test_forest_sets_pick_ray_all_and_dual_collision(gpu_ctx):
    """Sets of one forest answer like separately built trees: pick and ray-all with
    per-query set ids, and collision between two sets of the forest."""
    import octree_amd

    D = 3
    sizes = [3000, 1, 800, 5000]
    sets = [lognormal_boxes(n, D, 0.0, 500.0, np.log(2.0), 0.6, seed=90 + i) for i, n in enumerate(sizes)]
    f = octree_amd.StaticOctreeBox.forest(sets, 8, None, None, ctx=gpu_ctx)
    rng = np.random.default_rng(95)
    nq = 3000
    sid = rng.integers(0, len(sizes), nq).astype(np.uint32)
    picks = uniform_points(nq, D, 0.0, 500.0, seed=96)
    rays = np.concatenate([uniform_points(nq, D, 0.0, 500.0, seed=97), unit_dirs(nq, D, seed=98)], axis=1)
    poff, pids = f.pick_batch(picks, set_ids=sid).csr()
    aoff, aids = f.ray_all_batch(rays, set_ids=sid).csr()
    trees = []
    for s in range(len(sizes)):
        o = O.OracleTree(D, "box", sets[s], None, 0, 8)
        trees.append(o)
        m = np.nonzero(sid == s)[0]
        ooff, oids = o.pick(picks[m])
        for j, qi in enumerate(m):
            assert sorted(pids[poff[qi]:poff[qi + 1]].tolist()) == sorted(oids[ooff[j]:ooff[j + 1]].tolist())
        ooff, oids = o.ray_all(rays[m])
        for j, qi in enumerate(m):
            assert aids[aoff[qi]:aoff[qi + 1]].tolist() == oids[ooff[j]:ooff[j + 1]].tolist()
    got = f.collide_with(f, BASE_TOL, 0, 3).astype(np.int64)
    assert np.array_equal(got, trees[0].collide_with(trees[3]).astype(np.int64))


# ---- PlaneSearch / PlanePositiveSegmentation / FrustumCulling ---------------------------------------
def test_plane_kats(gpu_ctx):
    for c in kat()["planes"]:
        g = _tree(gpu_ctx, c["kind"], c["dim"], np.array(c["geom"], float), None, c["depth"], c["cap"])
        pl = c["planes"]
        if c["mode"] == 0:
            got = g.PlaneSearch(pl[0][0], pl[0][1:], None, c["tol"])
        else:
            got = g.FrustumCulling(pl, None, c["tol"])
        assert sorted(got) == sorted(c["expect"]), c["src"]


def _random_planes(n, D, lo, hi, seed):
    rng = np.random.default_rng(seed)
    nrm = unit_dirs(n, D, seed=seed + 1)
    pts = rng.uniform(lo, hi, (n, D))
    return np.concatenate([np.einsum("ij,ij->i", nrm, pts)[:, None], nrm], axis=1)


@pytest.mark.parametrize("kind,D,tol", [("point", 3, 0.5), ("box", 3, BASE_TOL), ("box", 2, 0.0), ("point", 6, 0.01),
                                        ("box", 4, 0.1)])
def test_planes_vs_oracle(gpu_ctx, kind, D, tol):
    n = 20000
    if kind == "point":
        geom = uniform_points(n, D, 0.0, 100.0, seed=110 + D)
    else:
        geom = lognormal_boxes(n, D, 0.0, 100.0, np.log(1.0), 0.6, seed=111 + D)
    g = _tree(gpu_ctx, kind, D, geom, None, 6, 8)
    o = O.OracleTree(D, kind, geom, None, 6, 8)
    pl = _random_planes(400, D, 0.0, 100.0, seed=112)
    pl[0, 0] = -1e6  # everything on the positive side
    for mode in (0, 1):
        goff, gids = g.planes_batch(pl, mode, tol).csr()
        ooff, oids = o.planes(pl, mode, tol, threads=8)
        _same_sets(goff, gids, ooff, oids)
    fr = _random_planes(4 * 300, D, 20.0, 80.0, seed=113).reshape(300, 4, D + 1)
    goff, gids = g.planes_batch(fr, 2, tol).csr()
    ooff, oids = o.planes(fr, 2, tol, threads=8)
    _same_sets(goff, gids, ooff, oids)
    assert goff[-1] > 0


def _sorted_segments(off, ids):
    seg = np.repeat(np.arange(len(off) - 1), np.diff(off.astype(np.int64)))
    return ids[np.lexsort((ids, seg))]


def test_planes_many_queries_one_lane_walk(gpu_ctx):
    """Launches of 65,536 or more plane queries walk one query per lane (fewer take a
    wavefront per query): both walks give the oracle's sets."""
    D, n = 3, 8000
    geom = uniform_points(n, D, 0.0, 100.0, seed=120)
    g = _tree(gpu_ctx, "point", D, geom, None, 5, 8)
    o = O.OracleTree(D, "point", geom, None, 5, 8)
    pl = _random_planes(70_000, D, 0.0, 100.0, seed=121)
    fr = _random_planes(4 * 70_000, D, 40.0, 60.0, seed=122).reshape(70_000, 4, D + 1)
    for qs, mode in ((pl, 0), (fr, 2)):
        goff, gids = g.planes_batch(qs, mode, 0.25).csr()
        ooff, oids = o.planes(qs, mode, 0.25, threads=8)
        assert np.array_equal(goff, ooff)
        assert np.array_equal(_sorted_segments(goff, gids), _sorted_segments(ooff, oids))
        assert goff[-1] > 0


@pytest.mark.parametrize("kind", ["point", "box"])
def test_planes_mid_batch_wavefront_walk(gpu_ctx, kind):
    """Between 4,097 and 65,535 plane queries one wavefront walks each query (up to
    4,096 take four wavefronts each, 65,536 and more one lane each): the oracle's sets."""
    D, n = 3, 6000
    geom = uniform_points(n, D, 0.0, 100.0, seed=130) if kind == "point" else \
        lognormal_boxes(n, D, 0.0, 100.0, np.log(1.0), 0.6, seed=130)
    g = _tree(gpu_ctx, kind, D, geom, None, 5, 8)
    o = O.OracleTree(D, kind, geom, None, 5, 8)
    pl = _random_planes(6000, D, 0.0, 100.0, seed=131)
    fr = _random_planes(4 * 5000, D, 30.0, 70.0, seed=132).reshape(5000, 4, D + 1)
    for qs, mode in ((pl, 0), (pl, 1), (fr, 2)):
        goff, gids = g.planes_batch(qs, mode, 0.25).csr()
        ooff, oids = o.planes(qs, mode, 0.25, threads=8)
        assert np.array_equal(goff, ooff)
        assert np.array_equal(_sorted_segments(goff, gids), _sorted_segments(ooff, oids))
        assert goff[-1] > 0
