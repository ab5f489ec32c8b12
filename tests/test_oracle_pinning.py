"""CPU tests: pin the oracle (oracle/otg_oracle.c) against the reference's own
known-answer vectors, closed forms and brute force.  No GPU needed."""
import math

import numpy as np
import pytest

from helpers import (BASE_TOL, DBL_MAX, MT19937, box_minmax, check_node_table_rules, collision_brute,
                     collision_two_brute, gaussian_mixture, issue38_boxes, kat, knn_box_brute, knn_brute,
                     lognormal_boxes, minmaxdist_textbook, pick_brute, plane_relation, range_brute_points, ray_hit_points,
                     uniform_points, unit_dirs)
from oracle import oracle as O


# ---- Morton / node keys ------------------------------------------------------
def test_morton_kats():
    for c in kat()["morton_encode"]:
        assert O.morton_encode(c["grid"], c["dim"]) == c["code"], c["src"]


def test_node_key_kats_1d():
    for c in kat()["node_key_1d"]:
        key = O.node_key(c["loc"], c["depth"], c["depth"], 1)
        assert key == c["key"], c["src"]
        # GetDepthID / RemoveSentinelBit (general.tests.cpp:483-494)
        assert (key.bit_length() - 1) // 1 == c["depth"]
        assert key - (1 << (key.bit_length() - 1)) == c["loc"]


@pytest.mark.parametrize("D", [2, 3, 4, 5, 6])
def test_morton_closed_form_and_decode(D):
    rng = np.random.default_rng(D)
    M = min(19, 63 // D) if D > 2 else 15
    for _ in range(200):
        g = rng.integers(0, 1 << M, D)
        code = O.morton_encode(g, D)
        expect = 0
        for i in range(M):
            for d in range(D):
                if (int(g[d]) >> i) & 1:
                    expect |= 1 << (i * D + d)
        assert code == expect
        # Decode(GetNodeID(code, depth)) keeps the top `depth` levels of each grid id
        depth = int(rng.integers(0, M + 1))
        key = O.node_key(code, depth, M, D)
        gd = O.decode(key, depth, M, D)
        mask = ~((1 << (M - depth)) - 1)
        assert gd == [int(x) & mask for x in g]


def test_mt19937_standard_value():
    # [rand.predef]: the 10000th invocation of a default-constructed mt19937 is 4123659995
    r = MT19937()
    for _ in range(9999):
        r()
    assert r() == 4123659995


def test_estimate_max_depth():
    # utils.h:613-622: n <= cap -> 2; else clamp(floor(log2(n/cap)/D), 2, max)
    assert O.estimate_max_depth(10, 16, 3) == 2
    assert O.estimate_max_depth(1_000_000, 16, 3) == int(math.log2(1_000_000 // 16) / 3)
    assert O.estimate_max_depth(10**12, 1, 3) == 13
    assert O.estimate_max_depth(2**63, 1, 3) == 21  # clamped to MAX_THEORETICAL_DEPTH_ID = 63 / 3
    assert O.estimate_max_depth(10**9, 1, 6) == 4
    assert O.estimate_max_depth(17, 16, 3) == 2


# ---- node table ------------------------------------------------------------------
@pytest.mark.parametrize("D,n,cap,M,seed", [(2, 300, 3, 6, 0), (3, 500, 4, 5, 1), (3, 1000, 16, 4, 2),
                                            (6, 400, 8, 3, 3), (3, 64, 1, 7, 4), (2, 1, 1, 3, 5), (3, 17, 16, 4, 6)])
def test_node_table_rules_points(D, n, cap, M, seed):
    pts = uniform_points(n, D, seed=seed)
    if n > 20:
        pts[: n // 10] = pts[n // 10: 2 * (n // 10)]  # duplicates
    t = O.OracleTree(D, "point", pts, None, M, cap)
    check_node_table_rules(t.export(), n, cap, D, False)


@pytest.mark.parametrize("D,n,cap,M,seed,sigma", [(2, 300, 3, 6, 0, 1.0), (3, 800, 5, 6, 1, 1.5), (3, 400, 2, 4, 2, 2.5),
                                                  (2, 50, 1, 8, 3, 2.0)])
def test_node_table_rules_boxes(D, n, cap, M, seed, sigma):
    b = lognormal_boxes(n, D, 0.0, 100.0, np.log(1.0), sigma, seed=seed)
    t = O.OracleTree(D, "box", b, None, M, cap)
    check_node_table_rules(t.export(), n, cap, D, True)


def test_rule_checker_catches_planted_mistakes():
    """A wrong depth, a wrong child range and a wrong leaf must each fail the rule check."""
    pts = uniform_points(400, 3, seed=11)
    t = O.OracleTree(3, "point", pts, None, 5, 4).export()
    check_node_table_rules(t, 400, 4, 3, False)
    internal = int(np.nonzero(t["ccount"])[0][1])
    for field, delta in (("depth", 1), ("ccount", -1), ("ebegin", 1)):
        bad = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in t.items()}
        bad[field][internal] = int(bad[field][internal]) + delta
        with pytest.raises(AssertionError):
            check_node_table_rules(bad, 400, 4, 3, False)


def test_minpoint_geometry_contains_points():
    """Every point lies in its leaf's MinPoint box; sizes halve per depth; children
    lie inside their parent (ot_static_linear_core.h:346-368, ot_base.h:121-134)."""
    pts = uniform_points(2000, 3, seed=5)
    t = O.OracleTree(3, "point", pts, [0, 0, 0, 100, 100, 100], 6, 8).export()
    size = 100.0
    for i in range(len(t["depth"])):
        assert np.all(t["size"][i] == size / 2 ** int(t["depth"][i]))
        if t["ccount"][i] == 0:
            ids = t["ids"][t["ebegin"][i]:t["ebegin"][i] + t["ecount"][i]]
            p = pts[ids]
            assert np.all(p >= t["min"][i] - 1e-9) and np.all(p <= t["min"][i] + t["size"][i] + 1e-9)
        for c in range(t["cbegin"][i], t["cbegin"][i] + t["ccount"][i]):
            assert np.all(t["min"][c] >= t["min"][i] - 1e-9)


# ---- kNN ----------------------------------------------------------------------------
def test_knn_kats_point_set_no1():
    k = kat()["knn_point_set_no1"]
    t = O.OracleTree(2, "point", k["points"], None, k["max_depth"], k["cap"])
    for c in k["cases"]:
        off, ids = t.knn([c["q"]], c["k"], c.get("max_within", DBL_MAX), c.get("tol", BASE_TOL))
        assert sorted(ids.tolist()) == sorted(c["expect"]), c["src"]


def test_knn_kats_issue9():
    for c in kat()["knn_issue9"]:
        pts = np.array(c["points"])
        t = O.OracleTree(c["dim"], "point", pts, c["box"], c["max_depth"], c["cap"])
        off, ids = t.knn([c["q"]], 1)
        d2 = ((pts - np.array(c["q"])) ** 2).sum(axis=1)
        assert ids.tolist()[0] == int(np.argmin(d2)), c["src"]


@pytest.mark.parametrize("D,n,k,cap,M", [(2, 1000, 4, 5, 10), (3, 3000, 8, 16, 6), (6, 2000, 16, 32, 4), (3, 500, 1, 2, 8)])
def test_knn_oracle_vs_brute(D, n, k, cap, M):
    pts = uniform_points(n, D, seed=D * 7 + k)
    pts[:50] = pts[50:100]  # exact duplicates -> equal distances
    qs = uniform_points(300, D, seed=99 + D)
    qs[:20] = pts[:20]
    t = O.OracleTree(D, "point", pts, None, M, cap)
    off, ids = t.knn(qs, k)
    got = O.lists(off, ids)
    amb = 0
    for qi in range(len(qs)):
        exp, a = knn_brute(qs[qi], pts, k)
        if a:
            amb += 1
            continue
        assert got[qi] == exp, qi
    assert amb == 0


def test_knn_oracle_max_within():
    pts = uniform_points(2000, 3, seed=3)
    qs = uniform_points(200, 3, seed=4)
    t = O.OracleTree(3, "point", pts, None, 6, 8)
    for mw in (0.0, 1.0, 3.0):
        off, ids = t.knn(qs, 8, mw)
        got = O.lists(off, ids)
        for qi in range(len(qs)):
            exp, a = knn_brute(qs[qi], pts, 8, mw)
            if not a:
                assert got[qi] == exp


# ---- box-tree kNN (MinMaxDistance) ---------------------------------------------------
def test_box_minmax_closed_forms():
    """The restated MinMaxDistance against the textbook MINMAXDIST for points
    outside the box, the plain point-box distance for the optimistic value, and
    the inside rule (0 and the smallest far-face distance)."""
    rng = np.random.default_rng(5)
    for D in (2, 3, 6):
        boxes = lognormal_boxes(400, D, 0.0, 10.0, seed=D)
        for _ in range(20):
            q = rng.uniform(-2.0, 12.0, D)
            dmin, dmm = box_minmax(q, boxes, 0.0)
            lo, hi = boxes[:, :D], boxes[:, D:]
            outside = np.any((q < lo) | (q > hi), axis=1)
            plain = np.sqrt((np.maximum(0.0, np.maximum(lo - q, q - hi)) ** 2).sum(axis=1))
            np.testing.assert_allclose(dmin, plain, rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(dmm[outside], minmaxdist_textbook(q, boxes[outside]), rtol=1e-9, atol=1e-12)
            far = np.sqrt(np.min(np.maximum((q - lo) ** 2, (hi - q) ** 2), axis=1))
            np.testing.assert_array_equal(dmin[~outside], 0.0)
            np.testing.assert_allclose(dmm[~outside], far[~outside], rtol=1e-12)
            assert np.all(dmin <= dmm * (1 + 1e-12) + 1e-12)


def test_box_knn_kats():
    for c in kat()["knn_box_2d"]:
        boxes = np.array(c["boxes"])
        t = O.OracleTree(2, "box", boxes, None, c["depth"], c["cap"])
        off, ids = t.knn([c["q"]], c["k"], DBL_MAX, c.get("tol", BASE_TOL))
        assert sorted(ids.tolist()) == sorted(c["expect"]), c["src"]
        exp, _ = knn_box_brute(np.array(c["q"]), boxes, c["k"], c.get("tol", BASE_TOL))
        assert sorted(exp) == sorted(c["expect"]), c["src"]


@pytest.mark.parametrize("D,k,cap,M", [(2, 1, 4, 6), (3, 4, 8, 5), (3, 16, 16, 6), (4, 8, 8, 4)])
def test_box_knn_oracle_vs_brute(D, k, cap, M):
    boxes = lognormal_boxes(2500, D, 0.0, 100.0, np.log(0.8), 0.7, seed=D * 11 + k)
    qs = uniform_points(200, D, seed=77 + D)
    t = O.OracleTree(D, "box", boxes, None, M, cap)
    off, ids = t.knn(qs, k)
    got = O.lists(off, ids)
    amb = 0
    for qi in range(len(qs)):
        exp, a = knn_box_brute(qs[qi], boxes, k)
        if a:
            amb += 1
            continue
        assert got[qi] == exp, qi
    assert amb == 0


def test_box_knn_planted_mistake_is_caught():
    """A pessimistic rule that drops the largest-difference term (plain sum of
    far-face distances) gives a different result set on these queries."""
    boxes = lognormal_boxes(1500, 3, 0.0, 50.0, np.log(1.0), 0.6, seed=3)
    qs = uniform_points(100, 3, 0.0, 50.0, seed=4)
    t = O.OracleTree(3, "box", boxes, None, 5, 8)
    off, ids = t.knn(qs, 4)
    got = O.lists(off, ids)
    differ = 0
    for qi, q in enumerate(qs):
        dmin, _ = box_minmax(q, boxes, BASE_TOL)
        wrong = np.sqrt((np.maximum((q - boxes[:, :3]) ** 2, (boxes[:, 3:] - q) ** 2)).sum(axis=1))
        F = max(BASE_TOL, np.sort(wrong)[3] * (1 + BASE_TOL))
        sel = np.nonzero(dmin < F)[0]
        differ += int(sel[np.lexsort((sel, dmin[sel]))].tolist() != got[qi])
    assert differ > 0


# ---- range --------------------------------------------------------------------------
@pytest.mark.parametrize("D", [2, 3, 6])
def test_range_oracle_vs_brute(D):
    pts = uniform_points(3000, D, seed=D)
    t = O.OracleTree(D, "point", pts, None, 5, 8)
    rng = np.random.default_rng(D + 1)
    c = rng.uniform(0, 100, (200, D))
    e = rng.uniform(1, 40, (200, 1))
    ranges = np.concatenate([c - e, c + e], axis=1)
    ranges[0] = np.concatenate([np.full(D, -1.0), np.full(D, 101.0)])  # root shortcut
    ranges[1] = np.concatenate([np.full(D, 5.0), np.full(D, 5.0)])     # zero volume
    off, ids = t.range(ranges)
    got = O.lists(off, ids)
    assert got[0] == list(range(3000))
    assert got[1] == []
    for qi in range(2, len(ranges)):
        assert sorted(got[qi]) == range_brute_points(ranges[qi], pts), qi


# ---- collision ---------------------------------------------------------------------
def test_collision_issue38_procedure():
    """general.tests.cpp:4728-4772: 10 rounds of 1000 loose 3D boxes (mt19937(0)),
    depth 10, cap 5; pairs normalised + sorted == brute force."""
    rng = MT19937(0)
    for _ in range(kat()["issue38_collision"]["rounds"]):
        boxes = issue38_boxes(rng)
        t = O.OracleTree(3, "box", boxes, None, 10, 5)
        got = t.collide()
        exp = collision_brute(boxes)
        assert np.array_equal(got.astype(np.int64), exp)


def test_collision_lognormal_vs_brute():
    boxes = lognormal_boxes(1500, 3, 0.0, 60.0, seed=8)
    t = O.OracleTree(3, "box", boxes, None, 5, 8)
    assert np.array_equal(t.collide().astype(np.int64), collision_brute(boxes))


# ---- ray first --------------------------------------------------------------------------
def test_ray_kats():
    for c in kat()["ray_first_2d"]:
        t = O.OracleTree(2, "box", c["boxes"], None, c["depth"], c["cap"])
        ray = c["o"] + c["d"]
        off, ids = t.ray_first([ray], c["tol"], c.get("inc", 0.0), c.get("max_dist", DBL_MAX))
        assert ids.tolist() == c["expect"], c["src"]


def _hit_all(ray, boxes, tol, inc):
    """Vectorised BoxHitTest (internal_geometry_module.h:631-712) over all boxes."""
    D = len(ray) // 2
    o, dv = np.array(ray[:D]), np.array(ray[D:])
    zero = np.abs(dv) < 2.0 * np.finfo(float).eps
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.where(zero, np.nan, 1.0 / np.where(zero, 1.0, dv))
    mind = boxes[:, :D] - o
    maxd = boxes[:, D:] - o
    en = np.full(len(boxes), -DBL_MAX)
    ex = np.full(len(boxes), DBL_MAX)
    for d in range(D):
        if zero[d]:
            continue
        p0, p1 = mind[:, d] * inv[d], maxd[:, d] * inv[d]
        a, b = (p1, p0) if dv[d] < 0 else (p0, p1)
        en = np.where(en < a, a, en)
        ex = np.where(b < ex, b, ex)
    et = np.maximum(0.0, ex) * inc + tol
    en = en - 0.5 * (np.maximum(0.0, en) * inc + tol)
    ex = ex + 0.5 * (np.maximum(0.0, ex) * inc + tol)
    hit = ~((en >= ex) | (ex <= 0.0))
    for d in range(D):
        if zero[d]:
            hit &= ~((mind[:, d] > et) | (maxd[:, d] < -et))
    return hit, np.maximum(0.0, en), ex


def test_ray_oracle_vs_brute():
    boxes = lognormal_boxes(3000, 3, 0.0, 100.0, np.log(1.0), 0.6, seed=12)
    t = O.OracleTree(3, "box", boxes, None, 6, 8)
    o = uniform_points(300, 3, seed=13)
    d = unit_dirs(300, 3, seed=14)
    rays = np.concatenate([o, d], axis=1)
    off, ids = t.ray_first(rays)
    got = O.lists(off, ids)
    for qi in range(len(rays)):
        hit, en, ex = _hit_all(rays[qi], boxes, BASE_TOL, 0.0)
        idx = np.nonzero(hit)[0]
        if len(idx) == 0:
            assert got[qi] == []
            continue
        M = float(np.min(ex[idx] + BASE_TOL))
        sel = idx[en[idx] <= M]
        exp = sel[np.lexsort((sel, -en[sel]))].tolist()
        assert got[qi] == exp, qi


# ---- PickSearch (ot_query.h:426-466) -------------------------------------------------
def test_pick_kats():
    for c in kat()["pick"]:
        if "boxes" in c:
            t = O.OracleTree(c["dim"], "box", c["boxes"], None, c["depth"], c["cap"])
        else:
            t = O.OracleTree(c["dim"], "point", c["points"], None, c["depth"], c["cap"])
        off, ids = t.pick([c["p"]], c.get("tol", BASE_TOL))
        assert sorted(ids.tolist()) == sorted(c["expect"]), c["src"]


@pytest.mark.parametrize("tol", [BASE_TOL, 0.0, 0.5])
def test_pick_boxes_vs_brute(tol):
    boxes = lognormal_boxes(3000, 3, 0.0, 100.0, np.log(3.0), 0.6, seed=21)
    t = O.OracleTree(3, "box", boxes, None, 6, 8)
    picks = uniform_points(300, 3, seed=22)
    picks[:100] = boxes[:100, :3]          # on a box corner (border case)
    picks[100:150] = boxes[100:150, 3:]    # on a max corner
    off, ids = t.pick(picks, tol)
    got = O.lists(off, ids)
    for qi in range(len(picks)):
        assert sorted(got[qi]) == pick_brute(picks[qi], boxes, True, tol), qi


def test_pick_points_vs_brute():
    pts = np.round(uniform_points(4000, 3, seed=23), 1)  # many exact duplicates
    t = O.OracleTree(3, "point", pts, None, 6, 8)
    picks = np.concatenate([pts[:200], uniform_points(100, 3, seed=24)])
    for tol in (BASE_TOL, 0.06, 0.0):
        off, ids = t.pick(picks, tol)
        got = O.lists(off, ids)
        for qi in range(len(picks)):
            assert sorted(got[qi]) == pick_brute(picks[qi], pts, False, tol), (tol, qi)


# ---- RayIntersectedAll (ot_query.h:1929-2012) --------------------------------------------
def test_ray_all_kats():
    for c in kat()["ray_all"]:
        t = O.OracleTree(c["dim"], "box", c["boxes"], None, c["depth"], c["cap"])
        off, ids = t.ray_all([c["o"] + c["d"]], c["tol"], c.get("inc", 0.0), c.get("max_dist", DBL_MAX))
        if c.get("ordered"):
            assert ids.tolist() == c["expect"], c["src"]
        else:
            assert sorted(ids.tolist()) == sorted(c["expect"]), c["src"]


@pytest.mark.parametrize("max_dist", [DBL_MAX, 0.0, 60.0])
def test_ray_all_boxes_vs_brute(max_dist):
    boxes = lognormal_boxes(3000, 3, 0.0, 100.0, np.log(1.0), 0.6, seed=31)
    t = O.OracleTree(3, "box", boxes, None, 6, 8)
    rays = np.concatenate([uniform_points(300, 3, seed=32), unit_dirs(300, 3, seed=33)], axis=1)
    rays[0, 3:] = [1.0, 0.0, 0.0]  # axis-parallel (NaN slabs)
    off, ids = t.ray_all(rays, BASE_TOL, 0.0, max_dist)
    got = O.lists(off, ids)
    for qi in range(len(rays)):
        hit, en, ex = _hit_all(rays[qi], boxes, BASE_TOL, 0.0)
        if max_dist != 0.0:
            hit &= en <= max_dist
        sel = np.nonzero(hit)[0]
        exp = sel[np.lexsort((sel, en[sel]))].tolist()
        assert got[qi] == exp, qi


def test_ray_all_points_vs_brute():
    pts = gaussian_mixture(4000, 3, 6, 0.0, 100.0, 2.0, 8.0, seed=34)
    t = O.OracleTree(3, "point", pts, None, 6, 8)
    rays = np.concatenate([uniform_points(200, 3, seed=35), unit_dirs(200, 3, seed=36)], axis=1)
    for tol in (0.05, 0.5):
        off, ids = t.ray_all(rays, tol, 0.0, DBL_MAX)
        got = O.lists(off, ids)
        for qi in range(len(rays)):
            hit, en, ex = ray_hit_points(rays[qi], pts, tol, 0.0)
            sel = np.nonzero(hit)[0]
            exp = sel[np.lexsort((sel, en[sel]))].tolist()
            assert got[qi] == exp, (tol, qi)


# ---- two-tree CollisionDetection (ot_query.h:1462-1484) --------------------------------------
def test_collision_two_trees_kats():
    for c in kat()["collision_two_trees"]:
        tl = O.OracleTree(2, "box", c["left"], None, c["depth"], c["cap"])
        tr = O.OracleTree(2, "box", c["right"], None, c["depth"], c["cap"])
        got = sorted(map(tuple, tl.collide_with(tr).tolist()))
        assert got == sorted(map(tuple, c["expect"])), c["src"]


@pytest.mark.parametrize("D", [2, 3])
def test_collision_two_trees_vs_brute(D):
    left = lognormal_boxes(1500, D, 0.0, 60.0, seed=41 + D)
    right = lognormal_boxes(1000, D, 10.0, 80.0, seed=51 + D)
    tl = O.OracleTree(D, "box", left, None, 5, 8)
    tr = O.OracleTree(D, "box", right, None, 6, 4)
    got = tl.collide_with(tr).astype(np.int64)
    assert len(got) > 0
    assert np.array_equal(got, collision_two_brute(left, right))
    # and the other way round: the same pairs with the sides swapped
    back = tr.collide_with(tl).astype(np.int64)
    assert np.array_equal(back[np.lexsort((back[:, 0], back[:, 1]))][:, ::-1], got)


# ---- PlaneSearch / PlanePositiveSegmentation / FrustumCulling (ot_query.h:612-822) ------------------
def test_plane_kats():
    for c in kat()["planes"]:
        t = O.OracleTree(c["dim"], c["kind"], c["geom"], None, c["depth"], c["cap"])
        off, ids = t.planes([c["planes"]], c["mode"], c["tol"])
        assert sorted(ids.tolist()) == sorted(c["expect"]), c["src"]


def _random_planes(n, D, lo, hi, seed):
    rng = np.random.default_rng(seed)
    nrm = unit_dirs(n, D, seed=seed + 1)
    pts = rng.uniform(lo, hi, (n, D))
    dist = np.einsum("ij,ij->i", nrm, pts)
    return np.concatenate([dist[:, None], nrm], axis=1)


@pytest.mark.parametrize("kind,D,tol", [("point", 3, 0.5), ("box", 3, BASE_TOL), ("box", 2, 0.0), ("point", 4, 1.0)])
def test_planes_vs_brute(kind, D, tol):
    if kind == "point":
        geom = uniform_points(3000, D, seed=101)
    else:
        geom = lognormal_boxes(3000, D, 0.0, 100.0, np.log(2.0), 0.6, seed=102)
    t = O.OracleTree(D, kind, geom, None, 6, 8)
    pl = _random_planes(120, D, 0.0, 100.0, seed=103)
    pl[0, 0] = -1e6  # every entity on the positive side: the positive-node shortcut at the root
    for mode in (0, 1):
        off, ids = t.planes(pl, mode, tol)
        got = O.lists(off, ids)
        for qi in range(len(pl)):
            rel = plane_relation(geom, kind == "box", pl[qi, 0], pl[qi, 1:], tol)
            exp = np.nonzero(rel == 0)[0] if mode == 0 else np.nonzero(rel != -1)[0]
            assert sorted(got[qi]) == exp.tolist(), (mode, qi)
    # frustums of 4 planes each
    fr = _random_planes(4 * 60, D, 20.0, 80.0, seed=104).reshape(60, 4, D + 1)
    off, ids = t.planes(fr, 2, tol)
    got = O.lists(off, ids)
    for qi in range(len(fr)):
        ok = np.ones(geom.shape[0], bool)
        for k in range(4):
            ok &= plane_relation(geom, kind == "box", fr[qi, k, 0], fr[qi, k, 1:], tol) != -1
        assert sorted(got[qi]) == np.nonzero(ok)[0].tolist(), qi


# ---- threaded build (the PAR path's parallel encode and sort) ------------------------
@pytest.mark.parametrize("kind,threads", [("point", 3), ("point", 8), ("box", 5)])
def test_threaded_build_equals_serial(kind, threads):
    """Encode + sort split over threads and merged gives the serial build's table
    exactly (ids break key ties, so the merged order is the serial order)."""
    D = 3
    geom = uniform_points(20000, D, seed=4) if kind == "point" else lognormal_boxes(20000, D, 0.0, 100.0,
                                                                                    np.log(1.0), 0.6, seed=4)
    if kind == "point":
        geom[:500] = geom[500:1000]
    a = O.OracleTree(D, kind, geom, None, 6, 8).export()
    b = O.OracleTree(D, kind, geom, None, 6, 8, threads=threads).export()
    for key in a:
        if isinstance(a[key], np.ndarray):
            assert np.array_equal(a[key], b[key]), key
