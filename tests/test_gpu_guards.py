"""Input guards of the query entry points: per-query set ids outside the tree's
sets, and batches whose result would not fit 32-bit list positions, are rejected
before any walk reads them or any buffer is sized from them."""
import numpy as np
import pytest

from helpers import uniform_points

pytestmark = pytest.mark.gpu


def test_set_ids_beyond_the_forest_are_rejected(gpu_ctx):
    import torch

    import octree_amd
    from octree_amd._abi import OTG_ERR_INVALID_ARG, OtgError

    sets = [uniform_points(500, 3, seed=s) for s in range(3)]
    f = octree_amd.StaticOctreePoint.forest(sets, 8, None, None, ctx=gpu_ctx)
    qs = uniform_points(100, 3, seed=9)
    good = np.arange(100, dtype=np.uint32) % 3
    r = f.knn_batch(qs, 4, set_ids=good)
    assert r.csr()[0][-1] == 400
    r.free()
    bad = good.copy()
    bad[57] = 3
    dq = torch.from_numpy(qs).to("cuda:0")
    dr = torch.from_numpy(np.concatenate([qs, qs + 5.0], axis=1)).to("cuda:0")
    dev_bad = torch.from_numpy(bad.view(np.int32)).to("cuda:0")
    calls = [lambda: f.knn_batch(qs, 4, set_ids=bad),
             lambda: f.range_batch(np.concatenate([qs, qs + 5.0], axis=1), set_ids=bad),
             lambda: f.pick_batch(qs, set_ids=bad),
             lambda: f.knn_batch(dq, 4, set_ids=dev_bad),
             lambda: f.range_batch(dr, set_ids=dev_bad),
             lambda: f.pick_batch(dq, set_ids=dev_bad)]
    for call in calls:
        with pytest.raises(OtgError) as e:
            call()
        assert e.value.status == OTG_ERR_INVALID_ARG
    # the forest still answers after the rejected calls
    r = f.knn_batch(torch.from_numpy(qs).to("cuda:0"), 4, set_ids=torch.from_numpy(good.view(np.int32)).to("cuda:0"))
    assert r.csr()[0][-1] == 400
    r.free()
    f.free()


def test_result_beyond_32_bit_positions_is_rejected(gpu_ctx):
    """4,097 ranges that contain the whole tree of 2^20 points would return
    4,097 * 2^20 > 2^32 ids: the count scan reports the overflow and the call fails
    with OTG_ERR_UNSUPPORTED instead of sizing buffers from a wrapped total."""
    import octree_amd
    from octree_amd._abi import OTG_ERR_UNSUPPORTED, OtgError

    n = 1 << 20
    t = octree_amd.StaticOctreePoint(ctx=gpu_ctx)
    t.Create(uniform_points(n, 3, seed=3), 6, [0, 0, 0, 100, 100, 100], 64)
    whole = np.tile([[-1.0, -1.0, -1.0, 101.0, 101.0, 101.0]], (4097, 1))
    with pytest.raises(OtgError) as e:
        t.range_batch(whole)
    assert e.value.status == OTG_ERR_UNSUPPORTED
    r = t.range_batch(whole[:3])
    off, ids = r.csr()
    r.free()
    assert off[-1] == 3 * n and np.array_equal(ids[:n], np.arange(n))
    t.free()
