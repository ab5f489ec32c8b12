"""otg_knn_noalloc (the allocation-free kNN entry point) called through the C ABI
against otg_knn on the same tree and queries: identical CSR offsets and ids, on
two calls in a row (the second with a smaller batch, so stale scratch contents
of the first would show), plus the limits it documents."""
import ctypes as C

import numpy as np
import pytest

from helpers import gaussian_mixture, uniform_points

pytestmark = pytest.mark.gpu


def _noalloc(tree, dq, dsets, k, cap_ids, torch, max_within=None, tol=None):
    from octree_amd import _abi
    from octree_amd import BASE_TOLERANCE, DBL_MAX

    nq = dq.shape[0]
    off = torch.empty(nq + 1, dtype=torch.int64, device="cuda:0")
    ids = torch.empty(max(cap_ids, 1), dtype=torch.int32, device="cuda:0")
    nids = C.c_uint64(0)
    st = _abi.lib().otg_knn_noalloc(tree.handle, nq, C.c_void_p(dq.data_ptr()),
                                    None if dsets is None else C.c_void_p(dsets.data_ptr()), k,
                                    DBL_MAX if max_within is None else max_within,
                                    BASE_TOLERANCE if tol is None else tol, C.c_void_p(off.data_ptr()),
                                    C.c_void_p(ids.data_ptr()), cap_ids, C.byref(nids))
    torch.cuda.synchronize()
    return st, off.cpu().numpy().view(np.uint64), ids.cpu().numpy().view(np.uint32)[:nids.value], nids.value


def test_noalloc_knn_equals_knn_twice(gpu_ctx):
    import torch

    import octree_amd
    from octree_amd import _abi

    D = 3
    sets = [gaussian_mixture(n, D, 8, 0.0, 1000.0, 2.0, 20.0, seed=60 + i) for i, n in enumerate([30000, 700, 5000])]
    box = np.tile([[0.0] * D + [1000.0] * D], (3, 1))
    f = octree_amd.StaticOctreePoint.forest(sets, 8, None, box, ctx=gpu_ctx)
    rng = np.random.default_rng(4)
    for nq, k in ((50_000, 16), (3_000, 5)):
        qs = uniform_points(nq, D, 0.0, 1000.0, seed=nq)
        qs[:50] = sets[0][:50]  # exact hits: distance 0
        sid = rng.integers(0, 3, nq).astype(np.uint32)
        dq = torch.from_numpy(qs).to("cuda:0")
        ds = torch.from_numpy(sid.view(np.int32)).to("cuda:0")
        want_off, want_ids = f.knn_batch(dq, k, set_ids=ds).csr()
        st, off, ids, n = _noalloc(f, dq, ds, k, int(want_off[-1]), torch)
        assert st == _abi.OTG_OK, _abi.lib().otg_ctx_last_error(gpu_ctx.handle)
        assert n == int(want_off[-1])
        assert np.array_equal(off, want_off) and np.array_equal(ids, want_ids), (nq, k)
    f.free()


def test_noalloc_knn_band_and_box_tree(gpu_ctx):
    """Equal distances longer than the list (the band pass, from scratch too) on a
    point tree, and box-tree kNN (MinMaxDistance)."""
    import torch

    import octree_amd
    from octree_amd import _abi

    D = 3
    pts = np.repeat(uniform_points(400, D, seed=8), 40, axis=0)  # 40 copies of each point
    t = octree_amd.StaticOctreePoint(ctx=gpu_ctx)
    t.Create(pts, 6, None, 8)
    qs = pts[::97].copy()
    dq = torch.from_numpy(qs).to("cuda:0")
    want_off, want_ids = t.knn_batch(dq, 4).csr()
    assert np.diff(want_off).max() >= 40
    st, off, ids, n = _noalloc(t, dq, None, 4, int(want_off[-1]), torch)
    assert st == _abi.OTG_OK
    assert np.array_equal(off, want_off) and np.array_equal(ids, want_ids)
    t.free()

    from helpers import lognormal_boxes

    b = lognormal_boxes(20000, D, 0.0, 100.0, np.log(0.5), 0.8, seed=3)
    bt = octree_amd.StaticOctreeBox(ctx=gpu_ctx)
    bt.Create(b, 6, None, 8)
    dq = torch.from_numpy(uniform_points(2000, D, seed=5)).to("cuda:0")
    want_off, want_ids = bt.knn_batch(dq, 8).csr()
    st, off, ids, n = _noalloc(bt, dq, None, 8, int(want_off[-1]), torch)
    assert st == _abi.OTG_OK
    assert np.array_equal(off, want_off) and np.array_equal(ids, want_ids)
    bt.free()


def test_noalloc_limits(gpu_ctx):
    import torch

    import octree_amd
    from octree_amd import _abi

    t = octree_amd.StaticOctreePoint(ctx=gpu_ctx)
    t.Create(uniform_points(10000, 3, seed=1), 6, None, 16)
    dq = torch.from_numpy(uniform_points(1000, 3, seed=2)).to("cuda:0")
    # a short ids buffer: OTG_ERR_CAPACITY, with the size needed
    st, off, ids, n = _noalloc(t, dq, None, 8, 100, torch)
    assert st == _abi.OTG_ERR_CAPACITY and n == 8000
    # more queries than the limit, k beyond 64, a set id beyond the tree
    big = torch.zeros(((1 << 18) + 1, 3), dtype=torch.float64, device="cuda:0")
    assert _noalloc(t, big, None, 8, 10, torch)[0] == _abi.OTG_ERR_INVALID_ARG
    assert _noalloc(t, dq, None, 65, 100000, torch)[0] == _abi.OTG_ERR_UNSUPPORTED
    bad = torch.ones(1000, dtype=torch.int32, device="cuda:0")
    assert _noalloc(t, dq, bad, 8, 100000, torch)[0] == _abi.OTG_ERR_INVALID_ARG
    # still answers after the rejected calls
    want_off, want_ids = t.knn_batch(dq, 8).csr()
    st, off, ids, n = _noalloc(t, dq, None, 8, 8000, torch)
    assert st == _abi.OTG_OK and np.array_equal(off, want_off) and np.array_equal(ids, want_ids)
    t.free()
