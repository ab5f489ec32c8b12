"""Serving layer (otg_server_*): resident trees, continuous batching, LRU eviction.

Determinism: a request's ids served inside a mix of requests (coalesced with
others on the same tree, split back per request, overlapped across workers)
equal the ids of the direct query on the same tree built alone, and those are
anchored to the oracle for one request of each kind."""
import threading

import numpy as np
import pytest

from helpers import BASE_TOL, DBL_MAX, gaussian_mixture, lognormal_boxes, uniform_points, unit_dirs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _direct(ctx, kind, D, geom, M, cap):
    import octree_amd

    cls = octree_amd.StaticOrthoTreePointND(D) if kind == "point" else octree_amd.StaticOrthoTreeBoxND(D)
    t = cls(ctx=ctx)
    t.Create(geom, M, None, cap)
    return t


def _requests(seed, D, kind, count):
    """Mixed request stream on one tree: (query kind, payload, params) with batch sizes 1..3000."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nq = int(np.exp(rng.uniform(0.0, np.log(3000.0))))
        c = rng.integers(0, 3)
        if kind == "box" and c == 0:
            c = 1
        if c == 0:
            out.append(("knn", uniform_points(nq, D, seed=seed * 1000 + i), dict(k=int(rng.integers(1, 33)))))
        elif c == 1:
            lo = uniform_points(nq, D, seed=seed * 1000 + i)
            ext = rng.uniform(0.5, 8.0, (nq, D))
            out.append(("range", np.concatenate([lo, lo + ext], axis=1), dict(mode=int(rng.integers(0, 2)))))
        else:
            o = uniform_points(nq, D, lo=-10.0, hi=110.0, seed=seed * 1000 + i)
            out.append(("ray", np.concatenate([o, unit_dirs(nq, D, seed=seed * 7 + i)], axis=1), {}))
    return out


def _direct_result(t, q, payload, params):
    if q == "knn":
        return t.knn_batch(payload, params["k"]).csr()
    if q == "range":
        return t.range_batch(payload, params["mode"]).csr()
    return t.ray_first_batch(payload).csr()


def _submit(srv, tid, q, payload, params):
    if q == "knn":
        return srv.submit_knn(tid, payload, params["k"])
    if q == "range":
        return srv.submit_range(tid, payload, params["mode"])
    return srv.submit_ray(tid, payload)


@pytest.mark.parametrize("batch_queries", [None, 1000])
def test_served_equals_direct_in_mix(gpu_ctx, batch_queries):
    """Requests from four client threads, batched with the default limit or with one of
    1,000 queries (many small batches; a larger request forms a batch alone)."""
    import octree_amd

    trees = {
        1: ("point", 3, gaussian_mixture(60000, 3, 12, 0.0, 100.0, 0.5, 6.0, seed=1), 8, 16),
        2: ("box", 3, lognormal_boxes(30000, 3, 0.0, 100.0, seed=2), 6, 12),
        3: ("point", 2, uniform_points(40000, 2, seed=3), 9, 8),
    }
    srv = octree_amd.Server(0, hbm_budget=8 << 30, workers=2, batch_queries=batch_queries)
    with pytest.raises(octree_amd.OtgError):
        srv.set_batch_limit(0)
    direct = {}
    for tid, (kind, D, geom, M, cap) in trees.items():
        srv.add_tree(tid, kind, D, geom, M, None, cap)
        direct[tid] = _direct(gpu_ctx, kind, D, geom, M, cap)
    work = []
    for tid, (kind, D, _, _, _) in trees.items():
        work += [(tid,) + r for r in _requests(tid, D, kind, 24)]
    order = np.random.default_rng(5).permutation(len(work))
    work = [work[i] for i in order]

    # submit from several host threads at once, as independent clients would
    ids = [None] * len(work)

    def client(lo, hi):
        for i in range(lo, hi):
            tid, q, payload, params = work[i]
            ids[i] = _submit(srv, tid, q, payload, params)

    chunks = np.linspace(0, len(work), 5).astype(int)
    th = [threading.Thread(target=client, args=(chunks[j], chunks[j + 1])) for j in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    for i, (tid, q, payload, params) in enumerate(work):
        res, lat, qms = srv.wait(ids[i])
        off, got = res.csr()
        eoff, exp = _direct_result(direct[tid], q, payload, params)
        assert np.array_equal(off, eoff) and np.array_equal(got, exp), (tid, q, params, len(payload))
        assert lat >= qms >= 0.0
    st = srv.stats()
    assert st["requests"] == len(work)
    assert st["launches"] <= st["requests"]
    # same request again, alone: identical ids
    tid, q, payload, params = work[0]
    res, _, _ = srv.wait(_submit(srv, tid, q, payload, params))
    off, got = res.csr()
    eoff, exp = _direct_result(direct[tid], q, payload, params)
    assert np.array_equal(off, eoff) and np.array_equal(got, exp)
    for t in direct.values():
        t.free()
    srv.close()


def test_served_vs_oracle():
    import octree_amd

    pts = uniform_points(20000, 3, seed=11)
    srv = octree_amd.Server(0, hbm_budget=2 << 30, workers=2)
    srv.add_tree(7, "point", 3, pts, 7, None, 16)
    o = O.OracleTree(3, "point", pts, None, 7, 16)
    qs = uniform_points(3000, 3, seed=12)
    lo = uniform_points(500, 3, seed=13)
    rg = np.concatenate([lo, lo + 6.0], axis=1)
    a = srv.submit_knn(7, qs, 8)
    b = srv.submit_range(7, rg)
    ra, _, _ = srv.wait(a)
    rb, _, _ = srv.wait(b)
    got, exp = ra.csr(), o.knn(qs, 8, DBL_MAX, BASE_TOL, threads=8)
    assert np.array_equal(got[0], exp[0]) and np.array_equal(got[1], exp[1])
    # range lists come back sorted by id (the reference returns its walk order; compared as sets)
    got, exp = rb.csr(), o.range(rg, 0, threads=8)
    assert np.array_equal(got[0], exp[0])
    for qi in range(len(rg)):
        assert got[1][got[0][qi]:got[0][qi + 1]].tolist() == sorted(exp[1][exp[0][qi]:exp[0][qi + 1]].tolist())
    srv.close()


def test_lru_eviction_under_budget():
    import octree_amd

    pts = [uniform_points(200000, 3, seed=20 + i) for i in range(3)]
    srv = octree_amd.Server(0, hbm_budget=1 << 30, workers=1)
    srv.add_tree(1, "point", 3, pts[0], 8, None, 16)
    one = srv.stats()["resident_bytes"]
    assert one > 0
    # admission reserves an estimate (36 B x 3 per point + 64 MiB) on top of what is resident:
    # this budget admits a second tree next to one resident tree, but not a third next to two
    srv.close()
    srv = octree_amd.Server(0, hbm_budget=int(one * 1.5) + (200000 * 36 * 3 + (64 << 20)), workers=1)
    srv.add_tree(1, "point", 3, pts[0], 8, None, 16)
    srv.add_tree(2, "point", 3, pts[1], 8, None, 16)
    r, _, _ = srv.wait(srv.submit_knn(1, pts[0][:10], 1))  # touch tree 1: tree 2 becomes LRU
    off, ids = r.csr()
    assert np.array_equal(ids, np.arange(10, dtype=np.uint32))
    srv.add_tree(3, "point", 3, pts[2], 8, None, 16)
    st = srv.stats()
    assert st["evictions"] >= 1
    assert srv.has_tree(3) and srv.has_tree(1) and not srv.has_tree(2)
    assert srv.tree_state(2) == 2  # registered, evicted
    # a request on the evicted tree rebuilds it (evicting the least recently used one);
    # its detail shows the cold tree and the time spent making it resident
    r, d = srv.wait_detail(srv.submit_knn(2, pts[1][:5], 1))
    assert np.array_equal(r.csr()[1], np.arange(5, dtype=np.uint32))
    assert d["cold"] == 1.0 and d["batch_requests"] >= 1 and d["resident_ms"] > 0.0
    assert abs(d["latency_ms"] - (d["queue_ms"] + d["service_ms"])) < 1e-6 and d["resident_ms"] <= d["service_ms"]
    r2, d2 = srv.wait_detail(srv.submit_knn(2, pts[1][:5], 1))
    assert d2["cold"] == 0.0 and np.array_equal(r2.csr()[1], np.arange(5, dtype=np.uint32))
    st = srv.stats()
    assert st["rebuilds"] == 1 and srv.has_tree(2) and st["resident_trees"] == 2
    with pytest.raises(octree_amd.OtgError):
        srv.submit_knn(99, pts[1][:4], 1)  # never registered
    srv.remove_tree(1)
    assert srv.tree_state(1) == 0
    assert srv.stats()["registered_trees"] == 2
    srv.close()


def test_served_mix_under_eviction(gpu_ctx):
    """A budget that holds about half of the trees: requests keep evicting and
    rebuilding trees while others run; every result still equals the direct query."""
    import octree_amd

    specs = {}
    for i in range(6):
        if i % 2 == 0:
            specs[i] = ("point", 3, gaussian_mixture(120000, 3, 6, 0.0, 100.0, 0.5, 5.0, seed=40 + i), 8, 16)
        else:
            specs[i] = ("box", 3, lognormal_boxes(60000, 3, 0.0, 100.0, seed=40 + i), 6, 16)
    probe = octree_amd.Server(0, hbm_budget=8 << 30, workers=1)
    for tid, (kind, D, geom, M, cap) in specs.items():
        probe.add_tree(tid, kind, D, geom, M, None, cap)
    working = probe.stats()["resident_bytes"]
    probe.close()
    # room for about half of the trees plus one admission reservation (the largest estimate)
    budget = working // 2 + max(120000 * 36 * 3, 60000 * 60 * 2) + (64 << 20)
    srv = octree_amd.Server(0, hbm_budget=budget, workers=2)
    direct = {}
    for tid, (kind, D, geom, M, cap) in specs.items():
        srv.add_tree(tid, kind, D, geom, M, None, cap)
        direct[tid] = _direct(gpu_ctx, kind, D, geom, M, cap)
    work = []
    for tid, (kind, D, _, _, _) in specs.items():
        work += [(tid,) + r for r in _requests(100 + tid, D, kind, 10)]
    order = np.random.default_rng(9).permutation(len(work))
    work = [work[i] for i in order]
    ids = [_submit(srv, tid, q, payload, params) for tid, q, payload, params in work]
    for i, (tid, q, payload, params) in enumerate(work):
        res, _, _ = srv.wait(ids[i])
        off, got = res.csr()
        eoff, exp = _direct_result(direct[tid], q, payload, params)
        assert np.array_equal(off, eoff) and np.array_equal(got, exp), (tid, q, params, len(payload))
    st = srv.stats()
    assert st["evictions"] > 0 and st["rebuilds"] > 0 and st["resident_bytes"] <= budget
    for t in direct.values():
        t.free()
    srv.close()
