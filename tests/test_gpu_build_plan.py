"""Build plans (otg_build_plan_*) and step plans (otg_step_plan_*): a point-forest build
captured as a HIP graph and replayed on new contents of the same device points.  Every
replay must equal a fresh otg_forest_build of the same contents (itself held to the
oracle in test_gpu_build and test_gpu_configs): node tables field by field for every
set, and the kNN / range answers.  New contents that need more nodes than the plan
holds fail with OTG_ERR_CAPACITY and leave the tree without nodes (queries answer
empty, nothing faults); the step plan (build, then a kNN batch and a range batch, as one
graph) equals a fresh build followed by eager kNN and range calls."""
import numpy as np
import pytest

from helpers import gaussian_mixture, uniform_points

pytestmark = pytest.mark.gpu


def _sets(seed, nsets, D, clustered, lo=3.0, hi=4.7):
    rng = np.random.default_rng(seed)
    sizes = np.round(10.0 ** rng.uniform(lo, hi, nsets)).astype(np.int64)
    if clustered:
        arrs = [gaussian_mixture(int(n), D, 6, 0.0, 100.0, 0.3, 4.0, seed=seed * 100 + i)
                for i, n in enumerate(sizes)]
    else:
        arrs = [uniform_points(int(n), D, seed=seed * 100 + i) for i, n in enumerate(sizes)]
    offs = np.zeros(nsets + 1, np.uint64)
    offs[1:] = np.cumsum(sizes)
    return arrs, offs


def _refill(seed, offs, D, clustered):
    """New contents of the same shape (set sizes) as offs."""
    nsets = len(offs) - 1
    sizes = np.diff(offs.astype(np.int64))
    if clustered:
        arrs = [gaussian_mixture(int(n), D, 6, 0.0, 100.0, 0.3, 4.0, seed=seed * 100 + i)
                for i, n in enumerate(sizes)]
    else:
        arrs = [uniform_points(int(n), D, seed=seed * 100 + i) for i, n in enumerate(sizes)]
    assert len(arrs) == nsets
    return np.concatenate(arrs)


def _same_tables(t, fresh, nsets):
    for s in range(nsets):
        a, b = t.export(s), fresh.export(s)
        assert a.keys() == b.keys()
        for key in a:
            assert np.array_equal(a[key], b[key]), f"set {s}: {key}"


def _queries(seed, nsets, D, nq=6000):
    rng = np.random.default_rng(seed)
    q = rng.uniform(0.0, 100.0, (nq, D))
    sid = rng.integers(0, nsets, nq).astype(np.uint32)
    e = rng.uniform(0.5, 8.0, (nq, 1))
    c = np.clip(rng.uniform(0.0, 100.0, (nq, D)), e / 2, 100.0 - e / 2)
    return q, sid, np.concatenate([c - e / 2, c + e / 2], axis=1)


@pytest.mark.parametrize("D,clustered,depths", [(3, True, None), (3, False, 6), (2, True, None), (6, True, 3)])
def test_build_plan_replays_equal_fresh_builds(gpu_ctx, D, clustered, depths):
    import torch

    import octree_amd

    print(f"build plan D={D} clustered={clustered} depths={depths}", flush=True)
    nsets = 10
    arrs, offs = _sets(7 + D, nsets, D, clustered)
    boxes = np.tile(np.array([0.0] * D + [100.0] * D), (nsets, 1))
    cap = 8
    pts = torch.from_numpy(np.concatenate(arrs)).to("cuda")
    torch.cuda.synchronize()
    plan = octree_amd.BuildPlan(pts, offs, cap, boxes, depths, ctx=gpu_ctx, dim=D)
    try:
        for rep in range(3):
            if rep:
                pts.copy_(torch.from_numpy(_refill(50 + rep, offs, D, clustered)))
                torch.cuda.synchronize()
                plan.run()
            fresh = octree_amd.StaticTree.forest(None, cap, depths, boxes, ctx=gpu_ctx, dim=D, kind="point",
                                                 device_geom=(pts, offs))
            try:
                _same_tables(plan.tree, fresh, nsets)
                q, sid, rq = _queries(90 + rep, nsets, D)
                for k in (8, 16):
                    a = plan.tree.knn_batch(q, k, set_ids=sid).csr()
                    b = fresh.knn_batch(q, k, set_ids=sid).csr()
                    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"kNN k={k} rep {rep}"
                a = plan.tree.range_batch(rq, set_ids=sid).csr()
                b = fresh.range_batch(rq, set_ids=sid).csr()
                assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"range rep {rep}"
            finally:
                fresh.free()
    finally:
        plan.free()


def test_build_plan_capacity_overflow(gpu_ctx):
    import torch

    import octree_amd

    n, D, cap = 200_000, 3, 16
    rng = np.random.default_rng(3)
    tight = 50.0 + rng.uniform(0.0, 1e-3, (n, D))  # one cell at every depth: a handful of nodes
    spread = uniform_points(n, D, seed=4)          # thousands of nodes
    offs = np.array([0, n], np.uint64)
    box = np.array([[0.0] * D + [100.0] * D])
    pts = torch.from_numpy(tight).to("cuda")
    torch.cuda.synchronize()
    plan = octree_amd.BuildPlan(pts, offs, cap, box, None, ctx=gpu_ctx, dim=D)
    try:
        pts.copy_(torch.from_numpy(spread))
        torch.cuda.synchronize()
        with pytest.raises(octree_amd.OtgError) as err:
            plan.run()
        assert err.value.status == octree_amd._abi.OTG_ERR_CAPACITY
        # the tree was left without nodes: queries answer empty and read nothing
        q = uniform_points(500, D, seed=5)
        off, ids = plan.tree.knn_batch(q, 4).csr()
        assert off[-1] == 0 and len(ids) == 0
        # the contents it was made for replay again
        pts.copy_(torch.from_numpy(tight))
        torch.cuda.synchronize()
        plan.run()
        fresh = octree_amd.StaticTree.forest(None, cap, None, box, ctx=gpu_ctx, dim=D, kind="point",
                                             device_geom=(pts, offs))
        try:
            _same_tables(plan.tree, fresh, 1)
            a, b = plan.tree.knn_batch(q, 4).csr(), fresh.knn_batch(q, 4).csr()
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        finally:
            fresh.free()
    finally:
        plan.free()


def test_step_plan_equals_build_then_queries(gpu_ctx):
    import torch

    import octree_amd

    D, nsets, cap, k, nq = 3, 12, 16, 16, 20_000
    arrs, offs = _sets(11, nsets, D, True)
    boxes = np.tile(np.array([0.0] * D + [100.0] * D), (nsets, 1))
    pts = torch.from_numpy(np.concatenate(arrs)).to("cuda")
    q = torch.zeros((nq, D), dtype=torch.float64, device="cuda")
    sid = torch.zeros(nq, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rq = torch.zeros((nq, 2 * D), dtype=torch.float64, device="cuda")
    hq0, hs0, hr0 = _queries(199, nsets, D, nq)  # contents of the kind the replays get
    q.copy_(torch.from_numpy(hq0))
    rq.copy_(torch.from_numpy(hr0))
    sid.copy_(torch.from_numpy(hs0.astype(np.int32)))
    torch.cuda.synchronize()
    plan = octree_amd.BuildPlan(pts, offs, cap, boxes, None, ctx=gpu_ctx, dim=D)
    try:
        kp = plan.tree.knn_plan(q, k, set_ids=sid)
        rp = plan.tree.range_plan(rq, set_ids=sid)
        step = plan.step(kp, rp)
        for rep in range(3):
            hq, hs, hr = _queries(200 + rep, nsets, D, nq)
            if rep:
                pts.copy_(torch.from_numpy(_refill(300 + rep, offs, D, True)))
            q.copy_(torch.from_numpy(hq))
            rq.copy_(torch.from_numpy(hr))
            sid.copy_(torch.from_numpy(hs.astype(np.int32)))
            torch.cuda.synchronize()
            step.run()
            got, got_r = kp.csr(), rp.csr()
            fresh = octree_amd.StaticTree.forest(None, cap, None, boxes, ctx=gpu_ctx, dim=D, kind="point",
                                                 device_geom=(pts, offs))
            try:
                want = fresh.knn_batch(hq, k, set_ids=hs).csr()
                assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), f"step rep {rep}"
                want_r = fresh.range_batch(hr, set_ids=hs).csr()
                assert np.array_equal(got_r[0], want_r[0]) and np.array_equal(got_r[1], want_r[1]), f"range rep {rep}"
                _same_tables(plan.tree, fresh, nsets)
            finally:
                fresh.free()
            # eager work on the same stream between replays
            plan.tree.range_batch(np.concatenate([hq[:64], hq[:64] + 1.0], axis=1), set_ids=hs[:64]).csr()
    finally:
        plan.free()
