"""kNN plans (otg_knn_plan_*): one batch captured as a HIP graph and replayed on new
contents of the same device buffers.  Every replay must equal the eager otg_knn call
on the same contents (itself held to the oracle in test_gpu_queries), across
replays with eager work between them, with set ids, the persistent walk, tolerance
bands longer than a list, and a set id outside the tree."""
import numpy as np
import pytest

from helpers import gaussian_mixture, uniform_points
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _eq(a, b, what):
    assert np.array_equal(a[0], b[0]), what
    assert np.array_equal(a[1], b[1]), what


@pytest.mark.parametrize("D,k,nq,persistent", [(3, 8, 40_000, 0), (3, 16, 40_000, 1), (2, 5, 3_000, 0),
                                               (6, 32, 20_000, 1)])
def test_plan_replays_equal_eager(gpu_ctx, D, k, nq, persistent):
    import torch

    import octree_amd

    print(f"plan D={D} k={k} nq={nq} persistent={persistent}", flush=True)
    pts = uniform_points(120_000, D, seed=D + k)
    t = octree_amd.StaticOrthoTreePointND(D)(ctx=gpu_ctx)
    t.Create(pts, None, None, 16)
    gpu_ctx.set_tuning(persistent=persistent)
    q = torch.zeros((nq, D), dtype=torch.float64, device="cuda")
    plan = t.knn_plan(q, k)
    try:
        for rep in range(4):
            host = uniform_points(nq, D, seed=100 * rep + D)
            if rep == 2:
                host[: nq // 2] = pts[: nq // 2]  # queries on entities: distance-0 ties
            q.copy_(torch.from_numpy(host))
            torch.cuda.synchronize()
            got = plan.run().csr()
            want = t.knn_batch(q, k).csr()
            _eq(got, want, f"replay {rep}")
            assert got[0][-1] == nq * k
            # eager work on the same stream between replays
            t.range_batch(np.concatenate([host[:64], host[:64] + 0.05], axis=1)).csr()
    finally:
        plan.free()
        gpu_ctx.set_tuning()
        t.free()


def test_plan_forest_set_ids(gpu_ctx):
    import torch

    import octree_amd

    rng = np.random.default_rng(5)
    sets = [gaussian_mixture(int(n), 3, 6, 0.0, 100.0, 1.0, 5.0, seed=20 + i)
            for i, n in enumerate(rng.integers(500, 20_000, 9))]
    f = octree_amd.StaticOctreePoint.forest(sets, 16, None, None, ctx=gpu_ctx)
    nq = 30_000
    q = torch.zeros((nq, 3), dtype=torch.float64, device="cuda")
    sid = torch.zeros(nq, dtype=torch.int32, device="cuda")
    plan = f.knn_plan(q, 10, set_ids=sid)
    try:
        for rep in range(3):
            s = rng.integers(0, len(sets), nq).astype(np.uint32)
            host = rng.uniform(0.0, 100.0, (nq, 3))
            q.copy_(torch.from_numpy(host))
            sid.copy_(torch.from_numpy(s.astype(np.int32)))
            torch.cuda.synchronize()
            got = plan.run().csr()
            _eq(got, f.knn_batch(host, 10, set_ids=s).csr(), f"forest replay {rep}")
        # an id outside the forest fails the run; the next good run succeeds
        sid[7] = len(sets)
        torch.cuda.synchronize()
        with pytest.raises(octree_amd.OtgError):
            plan.run()
        sid[7] = 0
        torch.cuda.synchronize()
        s = sid.cpu().numpy().astype(np.uint32)
        _eq(plan.run().csr(), f.knn_batch(q.cpu().numpy(), 10, set_ids=s).csr(), "after a bad id")
    finally:
        plan.free()
        f.free()


def test_plan_band_longer_than_list(gpu_ctx):
    """Duplicates at a query: the band outgrows the list, the run completes it outside
    the graph, and the next replay (no band) returns to the graph's own CSR."""
    import torch

    import octree_amd

    pts = uniform_points(6000, 3, seed=31)
    pts[:300] = pts[5000]
    t = octree_amd.StaticOctreePoint(ctx=gpu_ctx)
    t.Create(pts, 6, None, 8)
    o = O.OracleTree(3, "point", pts, None, 6, 8)
    band_q = np.concatenate([pts[5000:5001], uniform_points(500, 3, seed=32)])
    plain_q = uniform_points(501, 3, seed=33)
    q = torch.from_numpy(band_q.copy()).cuda()
    plan = t.knn_plan(q, 5)
    try:
        for host, band in ((band_q, True), (plain_q, False), (band_q, True)):
            q.copy_(torch.from_numpy(host))
            torch.cuda.synchronize()
            off, ids = plan.run().csr()
            ooff, oids = o.knn(host, 5, threads=8)
            assert np.array_equal(off, ooff) and np.array_equal(ids, oids)
            assert (off[1] - off[0] == 301) == band
    finally:
        plan.free()
        t.free()


def test_plan_box_tree_and_limits(gpu_ctx):
    import torch

    import octree_amd

    rng = np.random.default_rng(8)
    c = rng.uniform(0.0, 100.0, (30_000, 2))
    e = rng.uniform(0.1, 2.0, (30_000, 2))
    boxes = np.concatenate([c - e, c + e], axis=1)
    t = octree_amd.StaticQuadtreeBox(ctx=gpu_ctx)
    t.Create(boxes, None, None, 12)
    host = rng.uniform(0.0, 100.0, (20_000, 2))
    q = torch.from_numpy(host).cuda()
    plan = t.knn_plan(q, 7, max_within=3.0)
    try:
        _eq(plan.run().csr(), t.knn_batch(host, 7, max_within=3.0).csr(), "box plan")
        with pytest.raises(octree_amd.OtgError):
            t.knn_plan(q, 65)
        with pytest.raises(ValueError):
            t.knn_plan(host, 4)
    finally:
        plan.free()
        t.free()
