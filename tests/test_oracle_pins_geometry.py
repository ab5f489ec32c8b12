"""CPU tests: pin the oracle's box RangeSearch (both modes, with the root and
contained-subtree shortcuts), its MBR node geometry, its loose-box location
(level and centre cell) and EstimateMaxDepth against closed forms written from
the reference's definitions, independent of the oracle's C code.  Each pin also
shows that a plausible mistake in the restated rule fails it."""
import math

import numpy as np
import pytest

from helpers import BASE_TOL, lognormal_boxes, uniform_points
from oracle import oracle as O


# ---------------------------------------------------------------------------
# node-table helpers (plain Python over an exported table)
# ---------------------------------------------------------------------------
def subtree_ranges(t):
    """Entity range [a, b) of every node's subtree: its own segment, then its
    children's ranges in order (ot_static_linear_core.h:204-212 segments)."""
    nn = len(t["depth"])
    a = t["ebegin"].astype(np.int64)
    b = np.empty(nn, np.int64)
    for i in reversed(range(nn)):
        cc = int(t["ccount"][i])
        b[i] = (b[int(t["cbegin"][i]) + cc - 1] if cc else int(t["ebegin"][i]) + int(t["ecount"][i]))
    return a, b


# ---------------------------------------------------------------------------
# box RangeSearch (ot_query.h:476-583)
# ---------------------------------------------------------------------------
def range_contains_box(r, bmin, bmax, tol):
    """IGM::DoesRangeContainBox per axis (internal_geometry_module.h:190-199), all axes."""
    D = len(bmin)
    for d in range(D):
        rmin, rmax = r[d], r[D + d]
        if rmin - tol > bmin[d] or bmin[d] > rmax + tol:
            return False
        if rmin - tol > bmax[d] or bmax[d] > rmax + tol:
            return False
    return True


def inside_test(r, B):
    """AreBoxesOverlapped(range, e, e1_must_contain_e2 = true) (general.h:296-308): no tolerance."""
    D = B.shape[1] // 2
    ok = np.ones(B.shape[0], bool)
    for d in range(D):
        ok &= ~((r[d] > B[:, d]) | (B[:, d] > r[D + d]))
        ok &= ~((r[d] > B[:, D + d]) | (B[:, D + d] > r[D + d]))
    return ok


def overlap_test(r, B, tol):
    """GetBoxRelation(range, e) == Overlapped (general.h:262-289): every axis strictly
    overlapping by more than the tolerance."""
    D = B.shape[1] // 2
    ok = np.ones(B.shape[0], bool)
    for d in range(D):
        ok &= (r[d] + tol < B[:, D + d]) & (B[:, d] + tol < r[D + d])
    return ok


def range_brute_boxes(r, boxes, t, mode, tol=BASE_TOL, shortcut=True, inside_with_tol=False):
    """Result set of RangeSearch on a box tree, from its definition:
      * the root's MBR inside the range (with tolerance) -> every entity (:488-512);
      * a zero-volume range -> nothing (:515-520);
      * otherwise an entity is returned iff it passes the entity test (mode) or a
        node on its root path has its MBR inside the range with tolerance -- that
        node's whole subtree is copied untested (:549-568).  Every such node's
        ancestors overlap the range, so the depth-first walk reaches it.
    The keyword arguments plant mistakes (no subtree shortcut, a tolerance in the
    Inside test) for the tests that show the pin catches them."""
    D = boxes.shape[1] // 2
    n = boxes.shape[0]
    mins, sizes = t["min"], t["size"]
    if range_contains_box(r, mins[0], mins[0] + sizes[0], tol):
        return list(range(n))
    vol = 1.0
    for d in range(D):
        vol *= r[D + d] - r[d]
    if vol <= 0.0:
        return []
    if mode == "inside":
        if inside_with_tol:
            ok = np.ones(n, bool)
            for d in range(D):
                ok &= (r[d] - tol <= boxes[:, d]) & (boxes[:, D + d] <= r[D + d] + tol)
        else:
            ok = inside_test(r, boxes)
    else:
        ok = overlap_test(r, boxes, tol)
    res = set(np.nonzero(ok)[0].tolist())
    if shortcut:
        a, b = subtree_ranges(t)
        ids = t["ids"]
        for i in range(len(t["depth"])):
            if range_contains_box(r, mins[i], mins[i] + sizes[i], tol):
                res.update(int(x) for x in ids[a[i]:b[i]])
    return sorted(res)


def _shortcut_boxes(D, seed):
    """Lognormal boxes plus a tight group of eight small boxes inside one max-depth
    cell of the range's upper corner; four of them end one ulp above the range's
    maximum.  The group's leaf lies inside the range within the tolerance, so the
    contained-subtree shortcut returns all eight; the Inside test rejects four."""
    b = lognormal_boxes(300, D, 0.0, 100.0, np.log(0.6), 0.5, seed=seed)
    lo, hi = 40.0, 53.125  # 53.125 = 17 cells of 100 / 32 (max depth 5)
    g = np.empty((8, 2 * D))
    for k in range(8):
        g[k, :D] = 45.0 + 0.01 * k
        g[k, D:] = 45.2 + 0.01 * k
        g[k, 0] = hi - 0.2 + 0.01 * k
        g[k, D] = np.nextafter(hi, np.inf) if k < 4 else hi - 0.01
    b = np.concatenate([b, g], axis=0)
    r = np.concatenate([np.full(D, lo), np.full(D, hi)])
    return b, r


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("mode", ["inside", "overlap"])
def test_box_range_oracle_vs_brute(D, mode):
    boxes, r_short = _shortcut_boxes(D, seed=D)
    n = boxes.shape[0]
    t = O.OracleTree(D, "box", boxes, None, 5, 6)
    e = t.export()
    rng = np.random.default_rng(7 + D)
    c = rng.uniform(0, 100, (150, D))
    w = rng.uniform(0.5, 30, (150, D))
    ranges = np.concatenate([c - w, c + w], axis=1)
    ranges[0] = np.concatenate([e["min"][0] - 1.0, e["min"][0] + e["size"][0] + 1.0])  # root shortcut
    ranges[1] = np.concatenate([np.full(D, 5.0), np.full(D, 5.0)])                     # zero volume
    ranges[2] = r_short                                                                 # subtree shortcut
    ranges[3] = np.concatenate([c[3], c[3] + np.r_[0.0, np.full(D - 1, 3.0)]])          # flat in one axis
    off, ids = t.range(ranges, overlap=(mode == "overlap"))
    got = O.lists(off, ids)
    assert got[0] == list(range(n))
    assert got[1] == [] and got[3] == []
    for qi in range(len(ranges)):
        assert sorted(got[qi]) == range_brute_boxes(ranges[qi], boxes, e, mode), (mode, qi)


def test_box_range_pin_catches_planted_mistakes():
    """Without the subtree shortcut, or with a tolerance in the Inside test, the
    restated rule no longer matches the oracle on the shortcut query."""
    D = 3
    boxes, r = _shortcut_boxes(D, seed=3)
    t = O.OracleTree(D, "box", boxes, None, 5, 6)
    e = t.export()
    off, ids = t.range(r[None, :])
    got = sorted(ids.tolist())
    assert got == range_brute_boxes(r, boxes, e, "inside")
    # the shortcut adds boxes the entity test rejects
    assert got != range_brute_boxes(r, boxes, e, "inside", shortcut=False)
    # one leaf holding a box that starts half a tolerance below the range minimum and
    # a box far outside it (so the leaf is not inside the range): Inside rejects the
    # first box, an Inside test with tolerance would accept it
    near = np.array([[12.0 - BASE_TOL * 0.5, 13.0, 13.0, 14.0, 14.0, 14.0], [30.0, 30.0, 30.0, 40.0, 40.0, 40.0]])
    t2 = O.OracleTree(D, "box", near, None, 3, 4)
    e2 = t2.export()
    r2 = np.concatenate([np.full(D, 12.0), np.full(D, 20.0)])
    off2, ids2 = t2.range(r2[None, :])
    got2 = sorted(ids2.tolist())
    assert got2 == range_brute_boxes(r2, near, e2, "inside") == []
    assert range_brute_boxes(r2, near, e2, "inside", inside_with_tol=True) == [0]


def test_point_range_subtree_shortcut_vs_brute():
    """Point trees: entity test DoesBoxContainPoint with tolerance (general.h:225-243);
    a node's MinPoint box inside the range copies its subtree.  On cell-aligned
    ranges the shortcut takes whole cells whose points all pass anyway."""
    D = 3
    pts = uniform_points(4000, D, seed=21)
    t = O.OracleTree(D, "point", pts, [0, 0, 0, 100, 100, 100], 4, 8)
    e = t.export()
    ranges = np.array([[25.0, 25.0, 25.0, 50.0, 50.0, 50.0], [0.0, 0.0, 0.0, 100.0, 100.0, 50.0],
                       [12.5, 0.0, 37.5, 37.5, 62.5, 100.0]])
    off, ids = t.range(ranges)
    got = O.lists(off, ids)
    a, b = subtree_ranges(e)
    for qi, r in enumerate(ranges):
        ok = np.ones(len(pts), bool)
        for d in range(D):
            ok &= (r[d] - BASE_TOL < pts[:, d]) & (pts[:, d] < r[D + d] + BASE_TOL)
        res = set(np.nonzero(ok)[0].tolist())
        for i in range(len(e["depth"])):
            if range_contains_box(r, e["min"][i], e["min"][i] + e["size"][i], BASE_TOL):
                res.update(int(x) for x in e["ids"][a[i]:b[i]])
        assert sorted(got[qi]) == sorted(res), qi


# ---------------------------------------------------------------------------
# MBR node geometry (InitializeMinimalNodeGeometry, ot_static_linear_core.h:280-327)
# ---------------------------------------------------------------------------
def mbr_restated(t, boxes, children=True, store_max=False):
    """Bottom-up: a node's box unites its own entities' boxes and its children's
    boxes (min, min + size); stored as (min, max - min).  The keyword arguments plant
    mistakes (children ignored, the max stored instead of the size)."""
    D = boxes.shape[1] // 2
    nn = len(t["depth"])
    ids = t["ids"]
    mn = np.empty((nn, D))
    sz = np.empty((nn, D))
    order = np.argsort(-t["depth"].astype(np.int64), kind="stable")
    for i in order:
        lo = np.full(D, np.finfo(np.float64).max)
        hi = np.full(D, -np.finfo(np.float64).max)
        a, c = int(t["ebegin"][i]), int(t["ecount"][i])
        for j in ids[a:a + c]:
            for d in range(D):
                lo[d] = boxes[j, d] if boxes[j, d] < lo[d] else lo[d]
                hi[d] = boxes[j, D + d] if hi[d] < boxes[j, D + d] else hi[d]
        if children:
            for ch in range(int(t["cbegin"][i]), int(t["cbegin"][i]) + int(t["ccount"][i])):
                for d in range(D):
                    cmax = mn[ch, d] + sz[ch, d]
                    lo[d] = mn[ch, d] if mn[ch, d] < lo[d] else lo[d]
                    hi[d] = cmax if hi[d] < cmax else hi[d]
        mn[i] = lo
        sz[i] = hi if store_max else hi - lo  # the reference's Sub(max, min)
    return mn, sz


@pytest.mark.parametrize("D,seed,cap,M", [(2, 1, 3, 6), (3, 2, 5, 5), (3, 3, 2, 7)])
def test_mbr_is_tight_union_of_subtree(D, seed, cap, M):
    b = lognormal_boxes(1500, D, 0.0, 100.0, np.log(1.0), 0.8, seed=seed)
    t = O.OracleTree(D, "box", b, None, M, cap)
    e = t.export()
    mn, sz = mbr_restated(e, b)
    assert np.array_equal(e["min"].view(np.uint64), mn.view(np.uint64))
    assert np.array_equal(e["size"].view(np.uint64), sz.view(np.uint64))
    assert np.array_equal((e["min"] + e["size"]).view(np.uint64), (mn + sz).view(np.uint64))
    # the min corner is exactly the subtree's smallest entity min; the max corner is
    # the largest entity max up to the rounding of min + size at each level
    a, bb = subtree_ranges(e)
    for i in range(len(e["depth"])):
        sub = b[e["ids"][a[i]:bb[i]]]
        assert np.array_equal(e["min"][i], sub[:, :D].min(axis=0))
        true_max = sub[:, D:].max(axis=0)
        assert np.all(np.abs(e["min"][i] + e["size"][i] - true_max) <= 8 * np.spacing(np.abs(true_max) + 1.0))
    assert (e["ccount"] > 0).sum() > 10


def test_mbr_pin_catches_planted_mistakes():
    b = lognormal_boxes(800, 3, 0.0, 100.0, np.log(1.0), 0.8, seed=9)
    e = O.OracleTree(3, "box", b, None, 5, 3).export()
    for kw in ({"children": False}, {"store_max": True}):
        with np.errstate(over="ignore", invalid="ignore"):  # childless restatement of empty internal nodes
            mn, sz = mbr_restated(e, b, **kw)
        assert not (np.array_equal(e["min"], mn) and np.array_equal(e["size"], sz)), kw


# ---------------------------------------------------------------------------
# loose-box location (GetLooseBoxGridData<2.0>, si_grid.h:221-285; si_mortongrid.h:78-97)
# ---------------------------------------------------------------------------
def morton(g, D):
    code = 0
    for i in range(32):
        for d in range(D):
            if i * D + d < 64 and (int(g[d]) >> i) & 1:
                code |= 1 << (i * D + d)
    return code


def loose_location(box, smin, smax, M, wrong_level=False):
    """(code, depth) of one box: root (0, 0) when the box is not inside the space
    with tolerance; else centre cell gc = (u32)((centre - min) * raster) and level
    L = bit_width(ceil(max_d size_d * raster_d) - 1) (0 when the ceil is 0), the
    code with its low D*L bits cleared at depth M - L.  wrong_level plants
    L = bit_width(ceil(...)) (one level too coarse at exact powers of two)."""
    D = len(smin)
    for d in range(D):
        lo, hi = box[d], box[D + d]
        if smin[d] - BASE_TOL > lo or lo > smax[d] + BASE_TOL or smin[d] - BASE_TOL > hi or hi > smax[d] + BASE_TOL:
            return 0, 0
    res = float(1 << M)
    raster = [1.0 if (smax[d] - smin[d]) == 0 else res / (smax[d] - smin[d]) for d in range(D)]
    g = [int((((box[d] + box[D + d]) * 0.5) - smin[d]) * raster[d]) for d in range(D)]
    rel = 0.0
    for d in range(D):
        x = (box[D + d] - box[d]) * raster[d]
        rel = x if rel < x else rel
    c = int(math.ceil(rel))
    L = 0 if c == 0 else ((c if wrong_level else c - 1).bit_length())
    code = (morton(g, D) >> (D * L)) << (D * L)
    return code, M - L


def _level_boxes(D, seed):
    """Boxes whose largest relative size lands exactly on, just below and just above
    powers of two of the grid cell, zero-size boxes, and boxes outside the space."""
    rng = np.random.default_rng(seed)
    M = 6
    cell = 100.0 / (1 << M)
    out = []
    for L in range(0, 5):
        for f in (1.0, 0.999, 1.001, 0.5001):
            s = cell * (1 << L) * f
            for _ in range(6):
                c = rng.uniform(20.0, 80.0, D)
                h = np.full(D, s / 2)
                h[1:] *= rng.uniform(0.1, 1.0, D - 1)
                out.append(np.concatenate([c - h, c + h]))
    for _ in range(20):
        c = rng.uniform(0.0, 100.0, D)
        out.append(np.concatenate([c, c]))
    out.append(np.concatenate([np.full(D, -5.0), np.full(D, 3.0)]))     # outside: root
    out.append(np.concatenate([np.full(D, 99.0), np.full(D, 100.0 + 1e-9)]))
    out.append(np.concatenate([np.full(D, 0.0), np.full(D, 100.0)]))    # the whole space
    return np.array(out), M


@pytest.mark.parametrize("D", [2, 3])
def test_loose_location_closed_form(D):
    b, M = _level_boxes(D, seed=D)
    smin, smax = np.zeros(D), np.full(D, 100.0)
    t = O.OracleTree(D, "box", b, np.concatenate([smin, smax]), M, 4)
    e = t.export()
    code = np.empty(len(b), np.uint64)
    dep = np.empty(len(b), np.int64)
    code[e["ids"]] = e["sorted_code"]
    dep[e["ids"]] = e["sorted_depth"]
    for i, box in enumerate(b):
        c, dd = loose_location(box, smin, smax, M)
        assert (int(code[i]), int(dep[i])) == (c, dd), (i, box.tolist())
        # invariant: the level is the smallest L with ceil(rel) <= 2^L
        if dd > 0:
            L = M - dd
            raster = (1 << M) / 100.0
            rel = max((box[D + d] - box[d]) * raster for d in range(D))
            cl = math.ceil(rel)
            assert cl <= (1 << L) and (L == 0 or cl > (1 << (L - 1)))
    assert len(set(dep.tolist())) >= 4


def test_loose_location_pin_catches_planted_mistake():
    D = 3
    b, M = _level_boxes(D, seed=5)
    smin, smax = np.zeros(D), np.full(D, 100.0)
    e = O.OracleTree(D, "box", b, np.concatenate([smin, smax]), M, 4).export()
    dep = np.empty(len(b), np.int64)
    dep[e["ids"]] = e["sorted_depth"]
    wrong = np.array([loose_location(box, smin, smax, M, wrong_level=True)[1] for box in b])
    assert np.any(wrong != dep)


# ---------------------------------------------------------------------------
# EstimateMaxDepth (detail/utils.h:613-622)
# ---------------------------------------------------------------------------
# (n, cap, D, expected): expected worked out by hand from
#   n <= cap -> 2, else clamp(floor(log2(n / cap) / D), 2, MAX_THEORETICAL_DEPTH)
# with integer division n / cap, e.g. 1e6 / 64 = 15625 and log2(15625) = 13.93 -> 4.
ESTIMATE_CASES = [
    (10, 16, 3, 2),             # fewer entities than the capacity
    (16, 16, 3, 2),             # n == cap
    (1_000, 64, 3, 2),          # 15 leaves, log2 = 3.9 / 3 = 1.3 -> clamped to 2
    (1_000_000, 16, 3, 5),      # 62500 -> 15.93 / 3 = 5.31
    (1_000_000, 64, 3, 4),      # 15625 -> 13.93 / 3 = 4.64
    (1_000_000, 8, 3, 5),       # 125000 -> 16.93 / 3 = 5.64
    (4_000_000, 32, 3, 5),      # 125000 -> 5.64
    (8_000_000, 32, 6, 2),      # 250000 -> 17.93 / 6 = 2.99
    (1 << 30, 1, 3, 10),        # exactly 2^30 leaves -> 10
    ((1 << 30) - 1, 1, 3, 9),   # one short of 2^30: 29.999... / 3 -> 9
    (1 << 24, 1, 2, 12),        # 24 / 2
    (1 << 40, 1, 2, 15),        # 20, clamped to 31 / 2 = 15 for quadtrees
    (1 << 63, 1, 3, 21),        # 21 = 63 / 3
    (127, 2, 6, 2),             # 63 leaves -> 5.98 / 6 -> clamped to 2
    (33, 16, 3, 2),             # 2 leaves
]


@pytest.mark.parametrize("n,cap,D,expect", ESTIMATE_CASES)
def test_estimate_max_depth_hand_values(n, cap, D, expect):
    assert O.estimate_max_depth(n, cap, D) == expect


def test_estimate_max_depth_table_catches_planted_mistakes():
    def with_ceil(n, cap, D):
        if n <= cap:
            return 2
        return min(max(math.ceil(math.log2(n // cap) / D), 2), (31 if D <= 2 else 63) // D)

    def single_precision(n, cap, D):  # log2 of the leaf count rounded to float32
        if n <= cap:
            return 2
        x = float(np.log2(np.float32(n // cap)))
        return min(max(int(x / D), 2), (31 if D <= 2 else 63) // D)

    def no_lower_clamp(n, cap, D):
        if n <= cap:
            return 2
        return min(int(math.log2(n // cap) / D), (31 if D <= 2 else 63) // D)

    for wrong in (with_ceil, single_precision, no_lower_clamp):
        assert any(wrong(n, c, D) != e for n, c, D, e in ESTIMATE_CASES), wrong.__name__
