"""Range plans (otg_range_plan_*): one RangeSearch batch captured as a HIP graph and
replayed on new contents of the same device buffers.  Every replay must equal the eager
otg_range call on the same contents (itself held to the oracle in test_gpu_queries and
test_gpu_configs), for point and box trees in both modes, with set ids and with eager
work between replays; results that outgrow the plan fail with OTG_ERR_CAPACITY and the
next fitting run answers again; the step plan joins build, kNN and range in one graph."""
import numpy as np
import pytest

from helpers import gaussian_mixture, lognormal_boxes, uniform_points

pytestmark = pytest.mark.gpu


def _boxes(rng, nq, D, lo_e, hi_e, hi=100.0):
    e = rng.uniform(lo_e, hi_e, (nq, 1))
    c = np.clip(rng.uniform(0.0, hi, (nq, D)), e / 2, hi - e / 2)
    return np.concatenate([c - e / 2, c + e / 2], axis=1)


def _eq(a, b, what):
    assert np.array_equal(a[0], b[0]), what
    assert np.array_equal(a[1], b[1]), what


@pytest.mark.parametrize("D,kind,mode,nq", [(3, "point", 0, 30_000), (2, "point", 0, 5_000), (3, "box", 0, 20_000),
                                             (3, "box", 1, 20_000), (2, "box", 1, 20_000)])
def test_range_plan_replays_equal_eager(gpu_ctx, D, kind, mode, nq):
    import torch

    import octree_amd

    print(f"range plan D={D} {kind} mode={mode} nq={nq}", flush=True)
    if kind == "point":
        geom = gaussian_mixture(150_000, D, 8, 0.0, 100.0, 0.5, 6.0, seed=D)
    else:
        geom = lognormal_boxes(100_000, D, 0.0, 100.0, np.log(0.3), 0.5, seed=D)
    t = octree_amd.StaticTree(ctx=gpu_ctx, dim=D, kind=kind)
    t.Create(geom, None, None, 16)
    rng = np.random.default_rng(40 + D)

    def contents():
        host = _boxes(rng, nq, D, 0.5, 12.0)
        host[:50] = [[-1.0] * D + [101.0] * D]  # whole-tree boxes: the root shortcut
        host[50:60, D:] = host[50:60, :D]      # zero-volume boxes: empty
        return host

    q = torch.zeros((nq, 2 * D), dtype=torch.float64, device="cuda")
    q.copy_(torch.from_numpy(contents()))
    torch.cuda.synchronize()
    plan = t.range_plan(q, mode)
    try:
        for rep in range(4):
            host = contents()
            q.copy_(torch.from_numpy(host))
            torch.cuda.synchronize()
            got = plan.run().csr()
            _eq(got, t.range_batch(host, mode).csr(), f"replay {rep}")
            t.knn_batch(host[:64, :D], 3).csr()  # eager work between replays
    finally:
        plan.free()
        t.free()


def test_range_plan_forest_set_ids_and_capacity(gpu_ctx):
    import torch

    import octree_amd

    rng = np.random.default_rng(9)
    sets = [gaussian_mixture(int(n), 3, 6, 0.0, 100.0, 0.5, 5.0, seed=70 + i)
            for i, n in enumerate(rng.integers(2_000, 40_000, 7))]
    f = octree_amd.StaticOctreePoint.forest(sets, 8, None, None, ctx=gpu_ctx)
    nq = 25_000
    q = torch.zeros((nq, 6), dtype=torch.float64, device="cuda")
    sid = torch.zeros(nq, dtype=torch.int32, device="cuda")
    small = lambda: _boxes(rng, nq, 3, 0.2, 3.0)
    q.copy_(torch.from_numpy(small()))
    sid.copy_(torch.from_numpy(rng.integers(0, len(sets), nq).astype(np.int32)))
    torch.cuda.synchronize()
    plan = f.range_plan(q, set_ids=sid)
    try:
        for rep in range(3):
            host, s = small(), rng.integers(0, len(sets), nq).astype(np.uint32)
            q.copy_(torch.from_numpy(host))
            sid.copy_(torch.from_numpy(s.astype(np.int32)))
            torch.cuda.synchronize()
            _eq(plan.run().csr(), f.range_batch(host, set_ids=s).csr(), f"forest replay {rep}")
        # boxes holding whole sets: far more ids than the plan was made for
        q.copy_(torch.from_numpy(np.tile(np.array([[-1.0] * 3 + [101.0] * 3]), (nq, 1))))
        torch.cuda.synchronize()
        with pytest.raises(octree_amd.OtgError) as err:
            plan.run()
        assert err.value.status == octree_amd._abi.OTG_ERR_CAPACITY
        # a set id outside the forest fails the run; fitting contents answer again
        host = small()
        q.copy_(torch.from_numpy(host))
        sid[3] = len(sets)
        torch.cuda.synchronize()
        with pytest.raises(octree_amd.OtgError):
            plan.run()
        sid[3] = 0
        torch.cuda.synchronize()
        s = sid.cpu().numpy().astype(np.uint32)
        _eq(plan.run().csr(), f.range_batch(host, set_ids=s).csr(), "after the failed runs")
    finally:
        plan.free()
        f.free()


@pytest.mark.parametrize("records", [0, 2])
def test_range_plan_kept_records(gpu_ctx, records):
    """The count walks' records inside the graph: replays with many more capped queries
    than the plan has record slots for (those are filled by the fill walk), and with 2
    records per query (most in the overflow area, which runs out)."""
    import torch

    import octree_amd

    D, nq = 3, 20_000
    geom = gaussian_mixture(200_000, D, 6, 0.0, 100.0, 0.5, 6.0, seed=77)
    t = octree_amd.StaticTree(ctx=gpu_ctx, dim=D, kind="point")
    t.Create(geom, None, None, 8)
    rng = np.random.default_rng(78)

    def contents(nmed):
        host = _boxes(rng, nq, D, 0.2, 1.0)
        host[:50] = [[-1.0] * D + [101.0] * D]             # whole-tree boxes: the plan's id capacity
        host[50:50 + nmed] = _boxes(rng, nmed, D, 8.0, 14.0)  # capped walks
        return host

    q = torch.zeros((nq, 2 * D), dtype=torch.float64, device="cuda")
    q.copy_(torch.from_numpy(contents(10)))
    torch.cuda.synchronize()
    try:
        gpu_ctx.set_tuning(range_records=records)
        plan = t.range_plan(q, 0)
        try:
            for rep, nmed in enumerate((10, 2500, 40)):
                host = contents(nmed)
                q.copy_(torch.from_numpy(host))
                torch.cuda.synchronize()
                got = plan.run().csr()
                _eq(got, t.range_batch(host, 0).csr(), f"replay {rep} ({nmed} larger boxes)")
        finally:
            plan.free()
    finally:
        gpu_ctx.set_tuning()
        t.free()
