"""GPU parity at the benchmark's own configurations, at full size.

* batched uneven forest: the 256-set, 39.2M-point workload bench.py times, at the
  leaf capacity its autotune picks (64) and at 8: every set's node table, and
  2,000 kNN (k = 16) and 2,000 RangeSearch queries per set out of the full
  12.8M-query launches, against the oracle; the same build, kNN and range launches
  replayed as one graph (build plan + kNN plan + range plan) give the same tables and
  ids;
* OctreeBox: 4,000,000 boxes, depth 8, capacity 32: node table, the whole
  self-collision pair list, and 100,000 rays first-hit;
* TreePointND<6>: 8,000,000 points, depth 8, capacity 32: node table and a
  20,000-query sample of the 1,000,000-query kNN launch;
* served stream: the 64 resident trees of the served workload under an HBM
  budget below the working set, every request's result against the oracle.

Inputs come from bench.py's generators (same seeds).  The oracle runs on up to
16 host threads (builds of different trees concurrently, queries in chunks)."""
import concurrent.futures as cf

import numpy as np
import pytest

from helpers import compare_tables_fast, csr_sort_rows, csr_take, oracle_trees
from oracle import oracle as O

pytestmark = pytest.mark.gpu

THREADS = 16


def _torch():
    import torch

    return torch


@pytest.fixture(scope="module")
def batched():
    import bench

    return bench.make_workload(256, 50_000)


def _csr_equal(got, want, what):
    go, gi = got
    wo, wi = want
    assert np.array_equal(np.diff(go), np.diff(wo)), f"{what}: list lengths differ"
    assert np.array_equal(gi, wi), f"{what}: ids differ"


@pytest.mark.parametrize("cap", [64, 8])
def test_batched_forest_at_bench_caps(gpu_ctx, batched, cap):
    import octree_amd

    torch = _torch()
    W = batched
    offs = W["offsets"]
    nsets, nq = len(offs) - 1, W["nq"]
    dev = torch.device("cuda:0")
    d_pts = torch.from_numpy(W["points"]).to(dev)
    d_knn = torch.from_numpy(W["knn"]).to(dev)
    d_rng = torch.from_numpy(W["range"]).to(dev)
    d_sid = torch.from_numpy(W["set_ids"].view(np.int32)).to(dev)
    box = np.tile(W["box"], (nsets, 1))
    f = octree_amd.StaticOctreePoint.forest(None, cap, None, box, ctx=gpu_ctx, device_geom=(d_pts, offs))
    print(f"forest cap {cap}: {f.info().node_count} nodes", flush=True)
    rk = f.knn_batch(d_knn, 16, set_ids=d_sid)
    rr = f.range_batch(d_rng, set_ids=d_sid)
    koff, kids = rk.csr()
    roff, rids = rr.csr()
    rk.free()
    rr.free()

    trees = oracle_trees([(3, "point", W["points"][int(offs[s]):int(offs[s + 1])], W["box"], 0, cap)
                          for s in range(nsets)], THREADS)
    for s in range(nsets):
        compare_tables_fast(f.export(s), trees[s].export(), False)
    f.free()

    # the graph path: the same build and kNN launch replayed as one graph (build plan +
    # kNN plan joined as a step plan), twice
    plan = octree_amd.BuildPlan(d_pts, offs, cap, box, None, ctx=gpu_ctx, dim=3)
    try:
        kp = plan.tree.knn_plan(d_knn, 16, set_ids=d_sid)
        rp = plan.tree.range_plan(d_rng, set_ids=d_sid)
        step = plan.step(kp, rp)
        for rep in range(2):
            step.run()
            _csr_equal(kp.csr(), (koff, kids), f"graph kNN cap {cap}, replay {rep}")
            _csr_equal(rp.csr(), (roff, rids), f"graph range cap {cap}, replay {rep}")
        for s in range(nsets):
            compare_tables_fast(plan.tree.export(s), trees[s].export(), False)
    finally:
        plan.free()

    sample = 2000
    rows = np.concatenate([s * nq + np.arange(sample) for s in range(nsets)])

    def oracle_set(s):
        q = slice(s * nq, s * nq + sample)
        return trees[s].knn(W["knn"][q], 16), trees[s].range(W["range"][q])

    with cf.ThreadPoolExecutor(max_workers=THREADS) as ex:
        per_set = list(ex.map(oracle_set, range(nsets)))
    ko = np.concatenate([[0], np.cumsum([len(k[1]) for k, _ in per_set])])
    ok_off = np.concatenate([[0]] + [k[0][1:] + ko[i] for i, (k, _) in enumerate(per_set)])
    ok_ids = np.concatenate([k[1] for k, _ in per_set])
    _csr_equal(csr_take(koff, kids, rows), (ok_off, ok_ids), f"kNN cap {cap}")
    ro = np.concatenate([[0], np.cumsum([len(r[1]) for _, r in per_set])])
    or_off = np.concatenate([[0]] + [r[0][1:] + ro[i] for i, (_, r) in enumerate(per_set)])
    or_ids = np.concatenate([r[1] for _, r in per_set])
    g_off, g_ids = csr_take(roff, rids, rows)
    assert np.array_equal(np.diff(g_off), np.diff(or_off)), f"range cap {cap}: list lengths differ"
    # the GPU returns each list sorted by id, the reference in its walk order
    assert np.array_equal(g_ids, csr_sort_rows(or_off, or_ids)), f"range cap {cap}: ids differ"


def test_box_tree_4m(gpu_ctx):
    import bench
    import octree_amd

    torch = _torch()
    boxes, rays, _ = bench.make_box_4m()
    dev = torch.device("cuda:0")
    t = octree_amd.StaticOctreeBox(ctx=gpu_ctx)
    t.Create(torch.from_numpy(boxes).to(dev), 8, None, 32)
    o = O.OracleTree(3, "box", boxes, None, 8, 32)
    compare_tables_fast(t.export(), o.export(), True)
    print("box table equal", flush=True)
    pairs = t.collide()
    opairs = o.collide()
    assert pairs.shape == opairs.shape and np.array_equal(pairs, opairs), "collision pairs differ"
    print(f"{len(pairs)} collision pairs equal", flush=True)
    r = t.ray_first_batch(torch.from_numpy(rays).to(dev))
    got = r.csr()
    r.free()
    want = o.ray_first(rays, threads=THREADS)
    _csr_equal(got, want, "ray first")
    t.free()


def test_point6d_8m(gpu_ctx):
    import bench
    import octree_amd

    torch = _torch()
    pts, qs, box = bench.make_point6d_8m()
    dev = torch.device("cuda:0")
    t = octree_amd.StaticOrthoTreePointND(6)(ctx=gpu_ctx)
    t.Create(torch.from_numpy(pts).to(dev), 8, box, 32)
    o = O.OracleTree(6, "point", pts, box, 8, 32)
    compare_tables_fast(t.export(), o.export(), False)
    print("6-D table equal", flush=True)
    r = t.knn_batch(torch.from_numpy(qs).to(dev), 16)
    off, ids = r.csr()
    r.free()
    rows = np.concatenate([np.arange(10_000), np.random.default_rng(1).choice(len(qs), 10_000, replace=False)])
    want = o.knn(qs[rows], 16, threads=THREADS)
    _csr_equal(csr_take(off, ids, rows), want, "6-D kNN")
    t.free()


def test_served_64_trees():
    """The served workload's 64 trees under a budget of 60% of their resident size
    (evictions and rebuilds happen), 120 requests of the mixed stream submitted at
    once from the client; every result equals the oracle's on the same tree."""
    import bench
    import octree_amd

    trees, reqs = bench.make_served(120)
    srv = octree_amd.Server(0, hbm_budget=200 << 30, workers=2)
    for i, (kind, g, cap) in enumerate(trees):
        srv.add_tree(i, kind, 3, g, None, None, cap)
    working = srv.stats()["resident_bytes"]
    srv.close()
    srv = octree_amd.Server(0, hbm_budget=int(working * 0.6), workers=2)
    for i, (kind, g, cap) in enumerate(trees):
        srv.add_tree(i, kind, 3, g, None, None, cap)
    ids = []
    for tid, q, payload, k in reqs:
        if q == "knn":
            ids.append(srv.submit_knn(tid, payload, k))
        elif q == "range":
            ids.append(srv.submit_range(tid, payload))
        else:
            ids.append(srv.submit_ray(tid, payload))
    served = []
    for rid in ids:
        res, _, _ = srv.wait(rid)
        served.append(res.csr())
        res.free()
    st = srv.stats()
    srv.close()
    print(f"served 120 requests: {st['evictions']} evictions, {st['rebuilds']} rebuilds", flush=True)
    assert st["evictions"] > 0

    used = sorted({r[0] for r in reqs})
    otrees = dict(zip(used, oracle_trees([(3, trees[i][0], trees[i][1], None, 0, trees[i][2]) for i in used],
                                         THREADS)))
    for j, (tid, q, payload, k) in enumerate(reqs):
        o = otrees[tid]
        if q == "knn":
            _csr_equal(served[j], o.knn(payload, k, threads=THREADS), f"request {j} kNN")
        elif q == "range":
            wo, wi = o.range(payload, threads=THREADS)
            go, gi = served[j]
            assert np.array_equal(go, wo), f"request {j} range lengths"
            assert np.array_equal(gi, csr_sort_rows(wo, wi)), f"request {j} range ids"
        else:
            _csr_equal(served[j], o.ray_first(payload, threads=THREADS), f"request {j} ray")
