"""The library autotuner (otg_autotune): it picks the leaf capacity, the finer
index's capacity, the kNN walks' occupancy and the query order for a batched
workload, caches the pick per workload shape, and never changes an id: kNN and
range results under the tuned settings, and under forced settings of every knob,
equal the untuned ones."""
import numpy as np
import pytest

from helpers import gaussian_mixture, uniform_points

pytestmark = pytest.mark.gpu


def _workload(nsets=12, nq=30_000, seed=3):
    rng = np.random.default_rng(seed)
    sizes = np.round(10.0 ** rng.uniform(3.0, 4.7, nsets)).astype(np.int64)
    sets = [gaussian_mixture(int(n), 3, 8, 0.0, 1000.0, 2.0, 20.0, seed=seed * 100 + i) for i, n in enumerate(sizes)]
    offs = np.zeros(nsets + 1, np.uint64)
    offs[1:] = np.cumsum(sizes)
    pts = np.concatenate(sets)
    sid = rng.integers(0, nsets, nq).astype(np.uint32)
    knn = pts[rng.integers(0, len(pts), nq)] + rng.normal(0.0, 2.0, (nq, 3))
    c = rng.uniform(0.0, 1000.0, (nq, 3))
    e = rng.uniform(1.0, 30.0, (nq, 1))
    rq = np.concatenate([c - e / 2, c + e / 2], axis=1)
    box = np.tile([0.0, 0.0, 0.0, 1000.0, 1000.0, 1000.0], (nsets, 1))
    return sets, offs, pts, knn, sid, rq, box


def _run(ctx, sets, box, cap, knn, sid, rq):
    import octree_amd

    f = octree_amd.StaticOctreePoint.forest(sets, cap, None, box, ctx=ctx)
    k = f.knn_batch(knn, 16, set_ids=sid).csr()
    r = f.range_batch(rq, set_ids=sid).csr()
    f.free()
    return k, r


def test_tuned_results_equal_untuned():
    import octree_amd

    sets, offs, pts, knn, sid, rq, box = _workload()
    ctx = octree_amd.Context(0)
    try:
        base = _run(ctx, sets, box, 16, knn, sid, rq)
        tune = ctx.autotune("point", 3, offs, pts, box, None, knn, sid, 16, rq, sid)
        assert tune["leaf_capacity"] in (8, 16, 32, 64)
        assert tune["index_capacity"] in (0, 16, 32, 64)
        assert tune["waves_per_simd"] in (0, 1, 2) and tune["order_queries"] in (0, 1)
        assert tune["persistent"] in (0, 1)
        assert tune["step_ms"] > 0.0
        assert ctx.tuning()["waves_per_simd"] == tune["waves_per_simd"]
        got = _run(ctx, sets, box, tune["leaf_capacity"], knn, sid, rq)
        for g, b in zip(got, base):
            assert np.array_equal(g[0], b[0]) and np.array_equal(g[1], b[1])
        # the pick is cached for the shape: a second call returns it unchanged
        assert ctx.autotune("point", 3, offs, pts, box, None, knn, sid, 16, rq, sid) == tune
        # every knob forced away from its default leaves the ids alone
        for setting in (dict(index_capacity=8), dict(waves_per_simd=1), dict(order_queries=0), dict(persistent=1),
                        dict(range_records=2), dict(index_capacity=64, waves_per_simd=2, order_queries=0, persistent=1)):
            ctx.set_tuning(**setting)
            for cap in (8, 64):
                got = _run(ctx, sets, box, cap, knn, sid, rq)
                for g, b in zip(got, base):
                    assert np.array_equal(g[0], b[0]) and np.array_equal(g[1], b[1]), (setting, cap)
    finally:
        ctx.close()


def test_autotune_rejects_bad_arguments():
    import octree_amd
    from octree_amd._abi import OTG_ERR_INVALID_ARG, OtgError

    ctx = octree_amd.Context(0)
    try:
        with pytest.raises(OtgError) as e:
            ctx.set_tuning(waves_per_simd=9)
        assert e.value.status == OTG_ERR_INVALID_ARG
        pts = uniform_points(1000, 3, seed=1)
        with pytest.raises(OtgError) as e:  # kNN queries without k
            ctx.autotune("point", 3, [0, 1000], pts, None, None, pts[:10], None, 0)
        assert e.value.status == OTG_ERR_INVALID_ARG
    finally:
        ctx.close()
