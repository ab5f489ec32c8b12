"""Shared test helpers: seeded generators, std::mt19937, brute-force checkers,
and the structural rule checker of the static linear tree's node table.

The brute-force checkers restate each query's *result set* from its closed-form
definition without any tree, so they pin the oracle's and the GPU's traversals;
the rule checker validates a node table against the build rules of
core/ot_static_linear_core.h:417-519 without running a BFS.
"""
from __future__ import annotations

import json
import struct
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
EPS = float(np.finfo(np.float64).eps)
BASE_TOL = EPS * 100.0
DBL_MAX = float(np.finfo(np.float64).max)


def kat():
    return json.loads((GOLDEN / "kat_reference.json").read_text())


# ---------------------------------------------------------------------------
# std::mt19937 (C++ [rand.eng.mt], standardised output), for the reference's
# seeded procedures (Issue36 / Issue38).
# ---------------------------------------------------------------------------
class MT19937:
    def __init__(self, seed: int = 5489):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.idx = 624

    def _twist(self):
        mt = self.mt
        for i in range(624):
            y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
            v = mt[(i + 397) % 624] ^ (y >> 1)
            if y & 1:
                v ^= 0x9908B0DF
            mt[i] = v
        self.idx = 0

    def __call__(self) -> int:
        if self.idx >= 624:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def issue38_boxes(rng: MT19937, n=1000, space_max=8.0, size_max=2.0):
    """general.tests.cpp:4728-4772 box generator (3D)."""
    b = np.zeros((n, 6))
    b[0, :3] = 0.0
    b[0, 3:] = size_max
    b[1, :3] = space_max - size_max
    b[1, 3:] = space_max
    for i in range(2, n):
        for d in range(3):
            mn = min(float(rng() % 10000) / 10000.0 * space_max, space_max - 0.1)
            mx = min(mn + float(rng() % 10000) / 10000.0 * size_max, space_max)
            b[i, d] = mn
            b[i, 3 + d] = mx
    return b


# ---------------------------------------------------------------------------
# seeded generators (exact doubles, shared by oracle and GPU)
# ---------------------------------------------------------------------------
def uniform_points(n, dim, lo=0.0, hi=100.0, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, (n, dim))


def lognormal_boxes(n, dim, lo=0.0, hi=1000.0, mu=np.log(0.5), sigma=0.6, seed=0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(lo, hi, (n, dim))
    h = np.exp(rng.normal(mu, sigma, (n, dim)))
    return np.concatenate([c - h, c + h], axis=1)


def gaussian_mixture(n, dim, k, lo, hi, smin, smax, seed=0):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(lo, hi, (k, dim))
    sig = rng.uniform(smin, smax, k)
    lab = rng.integers(0, k, n)
    pts = centers[lab] + rng.normal(0.0, 1.0, (n, dim)) * sig[lab, None]
    bad = ((pts < lo) | (pts > hi)).any(axis=1)
    while bad.any():
        m = int(bad.sum())
        lab2 = rng.integers(0, k, m)
        pts[bad] = centers[lab2] + rng.normal(0.0, 1.0, (m, dim)) * sig[lab2, None]
        bad = ((pts < lo) | (pts > hi)).any(axis=1)
    return pts


def unit_dirs(n, dim, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, 1.0, (n, dim))
    return v / np.sqrt((v * v).sum(axis=1))[:, None]


# ---------------------------------------------------------------------------
# brute-force result sets
# ---------------------------------------------------------------------------
def distances(q, pts):
    d2 = np.zeros(pts.shape[0])
    for d in range(pts.shape[1]):
        v = q[d] - pts[:, d]
        d2 = d2 + v * v
    return np.sqrt(d2)


def knn_brute(q, pts, k, max_within=DBL_MAX, tol=BASE_TOL):
    """Closed-form kNN set of ot_query.h:1208-1247: {d < F_final}, ascending by
    (distance, id).  Returns (ids, ambiguous) where ambiguous flags a distance
    exactly at the final bound (order-dependent in the reference)."""
    dist = distances(q, pts)
    F0 = DBL_MAX if max_within == DBL_MAX else max(tol, max_within * (1.0 + tol))
    cand = np.nonzero(dist < F0)[0]
    if len(cand) < k:
        order = cand[np.lexsort((cand, dist[cand]))]
        return order.tolist(), False
    dk = np.sort(dist[cand])[k - 1]
    F = max(tol, dk * (1.0 + tol))
    lim = min(F, F0) if F0 != DBL_MAX else F
    sel = np.nonzero(dist < lim)[0]
    amb = bool(np.any(dist == lim)) or (F0 != DBL_MAX and F > F0 and bool(np.any((dist >= F0) & (dist < F))))
    order = sel[np.lexsort((sel, dist[sel]))]
    return order.tolist(), amb


def box_minmax(q, boxes, tol):
    """Optimistic and pessimistic point-box distances in the operation order of
    MinMaxDistance2 (adapters/general.h:159-210), vectorised over the boxes."""
    D = len(q)
    q = np.asarray(q, np.float64)
    n = boxes.shape[0]
    min2 = np.zeros(n)
    mm2 = np.zeros(n)
    far_in = np.full(n, DBL_MAX)
    largest = np.zeros(n)
    inside = np.ones(n, bool)
    for d in range(D):
        min_dist = q[d] - boxes[:, d]
        max_dist = boxes[:, D + d] - q[d]
        in_c = (-tol <= min_dist) & (-tol <= max_dist)
        inside &= in_c
        a, b = min_dist * min_dist, max_dist * max_dist
        lo, hi = np.where(b < a, b, a), np.where(b < a, a, b)
        far_in = np.where(inside & (hi < far_in), hi, far_in)
        min2 = np.where(in_c, min2, min2 + lo)
        largest = np.where(largest < hi - lo, hi - lo, largest)
        mm2 = mm2 + hi
    min2 = np.where(inside, 0.0, min2)
    mm2 = np.where(inside, far_in, mm2 - largest)
    return np.sqrt(min2), np.sqrt(mm2)


def minmaxdist_textbook(q, boxes):
    """MINMAXDIST of Roussopoulos, Kelley, Vincent (1995) for points outside the
    box: min over axes k of (near-face distance_k^2 + sum of far-face distances_i^2, i != k)."""
    D = len(q)
    lo = np.minimum((q - boxes[:, :D]) ** 2, (boxes[:, D:] - q) ** 2)
    hi = np.maximum((q - boxes[:, :D]) ** 2, (boxes[:, D:] - q) ** 2)
    tot = hi.sum(axis=1)
    return np.sqrt(np.min(tot[:, None] - hi + lo, axis=1))


def knn_box_brute(q, boxes, k, tol=BASE_TOL):
    """Closed-form box kNN set (ot_query.h:1058-1164 with MinMaxDistance): F_final =
    tol_upper(k-th smallest pessimistic distance); result = {optimistic < F_final}
    ascending by (optimistic distance, id).  Returns (ids, ambiguous)."""
    dmin, dmm = box_minmax(q, boxes, tol)
    F = DBL_MAX
    if len(boxes) >= k:
        x = np.sort(dmm)[k - 1]
        F = max(tol, x * (1.0 + tol))
    sel = np.nonzero(dmin < F)[0]
    order = sel[np.lexsort((sel, dmin[sel]))]
    return order.tolist(), bool(np.any(dmin == F))


def range_brute_points(r, pts, tol=BASE_TOL):
    D = pts.shape[1]
    ok = np.ones(pts.shape[0], bool)
    for d in range(D):
        if tol != 0.0:
            ok &= (r[d] - tol < pts[:, d]) & (pts[:, d] < r[D + d] + tol)
        else:
            ok &= (r[d] <= pts[:, d]) & (pts[:, d] <= r[D + d])
    return np.nonzero(ok)[0].tolist()


def pick_brute(p, geom, is_box, tol=BASE_TOL):
    """PickSearch's entity test over all entities: boxes DoesBoxContainPoint
    (general.h:225-243), points ArePointsEqual (general.h:218-221)."""
    D = len(p)
    if is_box:
        ok = np.ones(geom.shape[0], bool)
        for d in range(D):
            if tol != 0.0:
                ok &= (geom[:, d] - tol < p[d]) & (p[d] < geom[:, D + d] + tol)
            else:
                ok &= (geom[:, d] <= p[d]) & (p[d] <= geom[:, D + d])
        return np.nonzero(ok)[0].tolist()
    d2 = np.zeros(geom.shape[0])
    for d in range(D):
        x = geom[:, d] - p[d]
        d2 = d2 + x * x
    return np.nonzero(d2 <= tol * tol)[0].tolist()


def ray_hit_points(ray, pts, tol, inc):
    """RayHitTester::Hit(point) (internal_geometry_module.h:599-626) over all points:
    (hit, enter, exit)."""
    D = len(ray) // 2
    o, dv = np.asarray(ray[:D]), np.asarray(ray[D:])
    v = pts - o
    t = np.zeros(pts.shape[0])
    for d in range(D):
        t = t + v[:, d] * dv[d]
    len2 = np.zeros(pts.shape[0])
    for d in range(D):
        len2 = len2 + v[:, d] * v[:, d]
    dist2 = np.maximum(0.0, len2 - t * t)
    ct = np.maximum(0.0, t) * inc + tol
    hit = (dist2 <= ct * ct) & (t >= -ct)
    return hit, np.maximum(0.0, t - ct), t + ct


def plane_relation(geom, is_box, dist, normal, tol):
    """Entity-plane relation over all entities (-1 negative, 0 hit, 1 positive):
    boxes GetBoxPlaneRelationAD(min, max - min) (internal_geometry_module.h:254-280),
    points GetPointPlaneRelation (general.h:322-335), same operation order."""
    D = len(normal)
    if is_box:
        radius = np.full(geom.shape[0], tol)
        dot = np.zeros(geom.shape[0])
        half = [(geom[:, D + d] - geom[:, d]) * 0.5 for d in range(D)]
        for d in range(D):
            radius = radius + half[d] * abs(normal[d])
        for d in range(D):
            dot = dot + normal[d] * (geom[:, d] + half[d])
        cp = dot - dist
        return np.where(cp + radius < 0.0, -1, np.where(cp - radius > 0.0, 1, 0))
    proj = np.zeros(geom.shape[0])
    for d in range(D):
        proj = proj + normal[d] * geom[:, d]
    return np.where(proj < dist - tol, -1, np.where(proj > dist + tol, 1, 0))


def boxes_overlap_strict(a, B, tol):
    """general.h:262-289 GetBoxRelation(a, B[i]) == Overlapped, vectorised over B."""
    D = a.shape[0] // 2
    over = np.ones(B.shape[0], bool)
    for d in range(D):
        over &= (a[d] + tol < B[:, D + d]) & (B[:, d] + tol < a[D + d])
    return over


def collision_brute(boxes, tol=BASE_TOL):
    n = boxes.shape[0]
    pairs = []
    for i in range(n - 1):
        m = boxes_overlap_strict(boxes[i], boxes[i + 1:], tol)
        for j in np.nonzero(m)[0]:
            pairs.append((i, i + 1 + int(j)))
    return np.array(pairs, dtype=np.int64).reshape(-1, 2)


def collision_two_brute(left, right, tol=BASE_TOL):
    """All (left id, right id) with GetBoxRelation == Overlapped, sorted."""
    pairs = []
    for i in range(left.shape[0]):
        for j in np.nonzero(boxes_overlap_strict(left[i], right, tol))[0]:
            pairs.append((i, int(j)))
    return np.array(pairs, dtype=np.int64).reshape(-1, 2)


# ---------------------------------------------------------------------------
# node table comparison and rule checks
# ---------------------------------------------------------------------------
def compare_tables(gpu: dict, ora: dict, is_box: bool):
    """Exact equality of two node tables; entity storage compared within runs of
    equal location keys (the reference's std::sort leaves their order open)."""
    for name in ("depth", "ebegin", "ecount", "ccount", "key", "min", "size"):
        a, b = gpu[name], ora[name]
        assert a.shape == b.shape, (name, a.shape, b.shape)
        if a.dtype.kind == "f":
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), f"{name} differs at " \
                f"{np.argwhere(a.view(np.uint64) != b.view(np.uint64))[:5].tolist()}"
        else:
            assert np.array_equal(a, b), f"{name} differs at {np.nonzero(a != b)[0][:10].tolist()}"
    m = ora["ccount"] > 0
    assert np.array_equal(gpu["cbegin"][m], ora["cbegin"][m]), "cbegin differs"
    # entity storage: equal multiset per equal-key run
    if is_box:
        db = max(1, int(ora["max_depth"]).bit_length())
        okeys = (ora["sorted_code"].astype(np.uint64) << np.uint64(db)) | ora["sorted_depth"].astype(np.uint64)
    else:
        okeys = ora["sorted_code"]
    gkeys = gpu["sorted_keys"]
    assert np.array_equal(gkeys, okeys), "sorted location keys differ"
    n = len(okeys)
    if n:
        starts = np.concatenate([[0], np.nonzero(okeys[1:] != okeys[:-1])[0] + 1, [n]])
        ga, oa = gpu["ids"], ora["ids"]
        for s, e in zip(starts[:-1], starts[1:]):
            if e - s == 1:
                assert ga[s] == oa[s], ("entity id", s)
            else:
                assert sorted(ga[s:e]) == sorted(oa[s:e]), ("entity run", s, e)


def compare_tables_fast(gpu: dict, ora: dict, is_box: bool):
    """compare_tables without per-run Python loops (for tables of millions of
    entities): entity ids are compared after sorting each equal-key run by id."""
    for name in ("depth", "ebegin", "ecount", "ccount", "key", "min", "size"):
        a, b = gpu[name], ora[name]
        assert a.shape == b.shape, (name, a.shape, b.shape)
        if a.dtype.kind == "f":
            bad = np.nonzero((a.view(np.uint64) != b.view(np.uint64)).reshape(a.shape[0], -1).any(axis=1))[0]
        else:
            bad = np.nonzero(a != b)[0]
        assert len(bad) == 0, f"{name} differs at nodes {bad[:10].tolist()}"
    m = ora["ccount"] > 0
    assert np.array_equal(gpu["cbegin"][m], ora["cbegin"][m]), "cbegin differs"
    if is_box:
        db = max(1, int(ora["max_depth"]).bit_length())
        okeys = (ora["sorted_code"].astype(np.uint64) << np.uint64(db)) | ora["sorted_depth"].astype(np.uint64)
    else:
        okeys = ora["sorted_code"]
    assert np.array_equal(gpu["sorted_keys"], okeys), "sorted location keys differ"
    ga = gpu["ids"][np.lexsort((gpu["ids"], okeys))]
    oa = ora["ids"][np.lexsort((ora["ids"], okeys))]
    assert np.array_equal(ga, oa), "entity storage differs beyond the order of equal keys"


def csr_take(off, ids, rows):
    """Sub-CSR of the listed rows (in the given order)."""
    off = np.asarray(off, np.int64)
    rows = np.asarray(rows, np.int64)
    lens = off[rows + 1] - off[rows]
    out_off = np.zeros(len(rows) + 1, np.int64)
    np.cumsum(lens, out=out_off[1:])
    starts = np.repeat(off[rows] - out_off[:-1], lens)
    return out_off, ids[np.arange(out_off[-1], dtype=np.int64) + starts]


def csr_sort_rows(off, ids):
    """Every row's ids in ascending order (vectorised)."""
    off = np.asarray(off, np.int64)
    row = np.repeat(np.arange(len(off) - 1), np.diff(off))
    return ids[np.lexsort((ids, row))]


def oracle_trees(specs, threads=16):
    """Oracle trees built concurrently (the C builds run without the GIL).
    specs: (D, kind, geom, box, max_depth, cap) tuples."""
    import concurrent.futures as cf

    from oracle import oracle as O

    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda a: O.OracleTree(*a), specs))


def check_node_table_rules(t: dict, n: int, cap: int, D: int, is_box: bool):
    """Structural rules of the static linear tree (ot_static_linear_core.h:417-519),
    checked directly on a node table + sorted codes/depths:
      * node 0 is the root (depth 0, key 1); children blocks are contiguous and
        listed in parent (BFS) order: child blocks concatenated = 1..N-1;
      * a node's subtree covers a contiguous entity range; internal nodes keep
        the entities whose depth equals the node depth (boxes) at the front, the
        children cover the rest in order, grouped by the Morton digit at
        depth+1, one child per distinct digit;
      * non-root node depth = LCA depth of its range = min(common digits of the
        first and last code, depth of the first entity);
      * leaf iff count <= cap, depth >= maxDepth, or (boxes) all entities stay
        at the node depth; the node key is (1 << depth*D) | code >> (M-depth)*D."""
    M = int(t["max_depth"])
    codes = [int(c) for c in t["sorted_code"]]
    deps = [int(x) for x in t["sorted_depth"]] if is_box else [M] * n
    nn = len(t["depth"])
    if n == 0:
        assert nn == 0
        return
    assert int(t["depth"][0]) == 0 and int(t["key"][0]) == 1
    # BFS contiguity
    expect_next = 1
    for i in range(nn):
        cc = int(t["ccount"][i])
        if cc:
            assert int(t["cbegin"][i]) == expect_next, ("bfs order", i)
            expect_next += cc
    assert expect_next == nn
    # ranges
    rng_of = [None] * nn

    def rng(i):
        if rng_of[i] is None:
            a = int(t["ebegin"][i])
            if int(t["ccount"][i]) == 0:
                rng_of[i] = (a, a + int(t["ecount"][i]))
            else:
                cb, cc = int(t["cbegin"][i]), int(t["ccount"][i])
                rng_of[i] = (a, rng(cb + cc - 1)[1])
        return rng_of[i]

    for i in reversed(range(nn)):
        rng(i)
    assert rng_of[0] == (0, n)
    for i in range(nn):
        a, b = rng_of[i]
        dep = int(t["depth"][i])
        cnt = b - a
        cc = int(t["ccount"][i])
        if i > 0:
            x = codes[a] ^ codes[b - 1]
            lcp = M if x == 0 else M - (x.bit_length() + D - 1) // D
            assert dep == min(lcp, deps[a]), ("lca depth", i, dep, lcp, deps[a])
        sh = (M - dep) * D
        assert int(t["key"][i]) == ((1 << (dep * D)) | (codes[a] >> sh)), ("node key", i)
        all_nr = is_box and all(deps[j] == dep for j in range(a, b))
        if cnt <= cap or dep >= M or all_nr:
            assert cc == 0, ("should be leaf", i)
            assert int(t["ecount"][i]) == cnt
        else:
            assert cc > 0, ("should split", i)
            nr = int(t["ecount"][i])
            for j in range(a, a + nr):
                assert deps[j] == dep
            if is_box and a + nr < b:
                assert deps[a + nr] != dep
            if not is_box:
                assert nr == 0
            pos = a + nr
            prev_digit = -1
            sh1 = (M - dep - 1) * D
            for c in range(int(t["cbegin"][i]), int(t["cbegin"][i]) + cc):
                ca, cb2 = rng_of[c]
                assert ca == pos, ("children cover in order", i, c)
                digs = {(codes[j] >> sh1) & ((1 << D) - 1) for j in range(ca, cb2)}
                assert len(digs) == 1, ("one digit per child", i, c)
                dg = digs.pop()
                assert dg > prev_digit
                prev_digit = dg
                assert int(t["depth"][c]) > dep
                pos = cb2
            assert pos == b


# ---- hand-derived static-core archives -----------------------------------------
# Written field by field from serialization.h:359-386, the std::vector /
# std::array / std::variant serializers (serialization/stl/*.h) and
# BinaryArchive::operator& (serialization/binary_archive.h:85-175), for two
# trees small enough to build by hand (node tables checked against the oracle).

def _arr(*xs, cnt="<I"):
    return struct.pack(cnt, len(xs)) + struct.pack(f"<{len(xs)}d", *xs)


def hand_archive_point_2d() -> bytes:
    """Quadtree of (0.5,0.5), (3.5,0.5), (0.5,3.5) in [0,4]^2, depth 2, cap 1:
    root (depth 0, entity segment {0,0}, 3 children) and three one-point leaves
    at depth 2 (path-compressed) with minimum points (0,0), (3,0), (0,3).
    At most 4 nodes for 3 entities -> NodeStorage256 (u8 begins and segments)."""
    b = struct.pack("<I", 0) + struct.pack("<I", 0)           # core, base version
    b += struct.pack("<Q", 1) + struct.pack("<I", 2)          # m_maxElementNum, m_maxDepthID
    b += struct.pack("<I", 4) + _arr(4.0, 4.0) + _arr(2.0, 2.0) + _arr(1.0, 1.0) + _arr(0.5, 0.5)  # m_nodeSize
    b += struct.pack("<I", 0)                                 # variant index: NodeStorage256
    b += struct.pack("<I", 2) + bytes([1, 4])                 # nodeChildSegmentBegins
    b += struct.pack("<I", 4) + bytes([0, 0, 0, 1, 1, 1, 2, 1])  # nodeEntitySegment {begin, length}
    b += struct.pack("<I", 4) + bytes([0, 2, 2, 2])           # m_nodeDepthIDs
    b += struct.pack("<I", 3) + struct.pack("<3I", 0, 1, 2)   # m_entityStorage
    b += struct.pack("<I", 4) + _arr(0.0, 0.0) + _arr(0.0, 0.0) + _arr(3.0, 0.0) + _arr(0.0, 3.0)  # min points
    return b


def hand_archive_box_2d() -> bytes:
    """Loose quadtree of boxes [0,1]^2 and [3,4]^2, no box space (bounding box
    [0,4]^2), depth 2, cap 1: root with MBR (0,0)+(4,4) and two leaves at depth 2
    whose MBRs are the boxes.  NodeStorage256."""
    b = struct.pack("<I", 0) + struct.pack("<I", 0)
    b += struct.pack("<Q", 1) + struct.pack("<I", 2)
    b += _arr(0.0, 0.0) + _arr(4.0, 4.0)                      # m_nominalTreeBox {minPoint, size}
    b += struct.pack("<I", 0)
    b += struct.pack("<I", 2) + bytes([1, 3])
    b += struct.pack("<I", 3) + bytes([0, 0, 0, 1, 1, 1])
    b += struct.pack("<I", 3) + bytes([0, 2, 2])
    b += struct.pack("<I", 2) + struct.pack("<2I", 0, 1)
    b += struct.pack("<I", 3) + _arr(0.0, 0.0) + _arr(4.0, 4.0) + _arr(0.0, 0.0) + _arr(1.0, 1.0) \
        + _arr(3.0, 3.0) + _arr(1.0, 1.0)                     # MBRs {minPoint, size}
    return b


HAND_POINTS_2D = [[0.5, 0.5], [3.5, 0.5], [0.5, 3.5]]
HAND_BOXES_2D = [[0.0, 0.0, 1.0, 1.0], [3.0, 3.0, 4.0, 4.0]]


def archive_children(begins, nn):
    """GetNodeChildren on an archived nodeChildSegmentBegins (ot_static_linear_core.h:191-202)."""
    out = []
    for i in range(nn):
        if i >= len(begins) - 1:
            out.append((0, 0))
        else:
            out.append((int(begins[i]), int(begins[i + 1])))
    return out


def check_archive_vs_table(a: dict, t: dict, is_box: bool, cap: int, space_min, space_max):
    """Archive fields against an oracle node table: header, children through the
    reference's GetNodeChildren rule, segments, depths, geometry bit for bit,
    and the entity storage as equal multisets per run of equal location keys."""
    D = t["min"].shape[1]
    nn = len(t["depth"])
    assert a["max_elements"] == cap and a["max_depth"] == t["max_depth"]
    size = np.asarray(space_max, float) - np.asarray(space_min, float)
    if is_box:
        assert np.array_equal(a["tree_box"][0], np.asarray(space_min, float))
        assert np.array_equal(a["tree_box"][1], size)
        assert np.array_equal(a["node_size_mbr"].view(np.uint64), t["size"].view(np.uint64))
    else:
        M = t["max_depth"]
        expect = np.stack([size * 0.5 ** k for k in range(M + 2)])
        assert np.array_equal(a["node_size"], expect)
    assert np.array_equal(a["depth"], t["depth"])
    assert np.array_equal(a["entity_begin"], t["ebegin"].astype(np.uint64))
    assert np.array_equal(a["entity_count"], t["ecount"].astype(np.uint64))
    assert np.array_equal(a["node_min"].view(np.uint64), t["min"].view(np.uint64))
    kids = archive_children(a["child_begins"], nn)
    for i in range(nn):
        lo, hi = kids[i]
        assert hi - lo == t["ccount"][i], ("children", i)
        if t["ccount"][i]:
            assert lo == t["cbegin"][i], ("first child", i)
    gi, oi = a["entity_ids"], t["ids"]
    assert len(gi) == len(oi)
    if is_box:
        db = max(1, int(t["max_depth"]).bit_length())
        keys = (t["sorted_code"].astype(np.uint64) << np.uint64(db)) | t["sorted_depth"].astype(np.uint64)
    else:
        keys = t["sorted_code"]
    n = len(keys)
    if n:
        starts = np.concatenate([[0], np.nonzero(keys[1:] != keys[:-1])[0] + 1, [n]])
        for s, e in zip(starts[:-1], starts[1:]):
            assert sorted(gi[s:e]) == sorted(oi[s:e]), ("entity run", s, e)
