"""octree_amd -- Python host mirror of OrthoTree's static-tree Core API over libotg.

The product is the C ABI in include/otg/otg.h (and the C++20 drop-in
include/otg/orthotree_gpu.hpp).  This module mirrors the reference's class and
method names (core/ot_aliases.h:82-91, :225-245; core/ot_query.h) for Python
callers, tests and the benchmark.  Every call runs on the GPU through libotg;
there is no CPU fallback.

Geometry is numpy float64: points (n, D); boxes (n, 2, D) or (n, 2*D) as
[min..., max...]; rays (n, 2*D) as [origin..., direction...].
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Optional, Sequence

import numpy as np

from . import _abi
from ._abi import OtgError

BASE_TOLERANCE = float(np.finfo(np.float64).eps * 100.0)  # adapters/general.h:57
DBL_MAX = float(np.finfo(np.float64).max)

PAR_EXEC = "PAR_EXEC"  # types.h:117-119 execution tags (the GPU path is the parallel one)
SEQ_EXEC = "SEQ_EXEC"
GPU_EXEC = PAR_EXEC


def _check(status: int, what: str, ctx: Optional["Context"] = None) -> None:
    if status != _abi.OTG_OK:
        detail = ""
        if ctx is not None and ctx.handle:
            detail = (_abi.lib().otg_ctx_last_error(ctx.handle) or b"").decode()
        raise OtgError(status, what, detail)


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else C.c_void_p(a.ctypes.data)


class Context:
    """One HIP device + stream (otg_ctx)."""

    def __init__(self, device: int = 0):
        L = _abi.lib()
        h = C.c_void_p()
        st = L.otg_ctx_create(device, C.byref(h))
        if st != _abi.OTG_OK:
            raise OtgError(st, "otg_ctx_create")
        self.handle = h
        self._children = weakref.WeakSet()  # trees / results that must be freed before the context

    def synchronize(self) -> None:
        _check(_abi.lib().otg_ctx_synchronize(self.handle), "otg_ctx_synchronize", self)

    def set_profiling(self, on: bool) -> None:
        _check(_abi.lib().otg_ctx_set_profiling(self.handle, 1 if on else 0), "set_profiling", self)

    def kernel_times(self) -> dict:
        L = _abi.lib()
        cap = 128
        names = C.create_string_buffer(cap * 64)
        ms = (C.c_double * cap)()
        nl = (C.c_uint64 * cap)()
        cnt = C.c_uint32()
        _check(L.otg_ctx_kernel_times(self.handle, cap, names, ms, nl, C.byref(cnt)), "kernel_times", self)
        out = {}
        for i in range(min(cnt.value, cap)):
            nm = names.raw[i * 64:(i + 1) * 64].split(b"\0", 1)[0].decode()
            out[nm] = (ms[i], nl[i])
        return out

    def reset_kernel_times(self) -> None:
        _check(_abi.lib().otg_ctx_reset_kernel_times(self.handle), "reset_kernel_times", self)

    def autotune(self, kind: str, dim: int, offsets, geom, boxes=None, max_depths=None, knn=None, knn_sets=None,
                 k: int = 0, ranges=None, range_sets=None) -> dict:
        """otg_autotune: measure the forest build of `geom` (device tensor or host array
        of all sets, split by `offsets`) and the kNN / range batches under candidate
        settings; the pick is applied to this context and returned (leaf_capacity is
        for the caller's builds)."""
        offs = np.ascontiguousarray(np.asarray(offsets, dtype=np.uint64))
        ns = len(offs) - 1
        cols = dim if kind == "point" else 2 * dim
        keep = []

        def arg(a, c, dt=np.float64):
            if a is None:
                return None, 0
            if _is_device(a):
                keep.append(a)
                return C.c_void_p(a.data_ptr()), a.shape[0]
            x = np.ascontiguousarray(np.asarray(a, dtype=dt).reshape((-1, c) if c else -1))
            keep.append(x)
            return C.c_void_p(x.ctypes.data), x.shape[0]

        gp, _ = arg(geom, cols)
        gmem = _abi.OTG_MEM_DEVICE if _is_device(geom) else _abi.OTG_MEM_HOST
        qp, nk = arg(knn, dim)
        sp_, _ = arg(knn_sets, 0, np.uint32)
        rp, nr = arg(ranges, 2 * dim)
        rsp, _ = arg(range_sets, 0, np.uint32)
        qdev = any(_is_device(a) for a in (knn, ranges) if a is not None)
        bx, _ = arg(boxes, 2 * dim)
        dp, _ = arg(max_depths, 0, np.uint32)
        out = _abi.Tuning()
        kindc = _abi.OTG_KIND_POINT if kind == "point" else _abi.OTG_KIND_BOX
        st = _abi.lib().otg_autotune(self.handle, kindc, dim, ns, offs.ctypes.data_as(_abi.U64P), gp, gmem, bx, dp,
                                     nk, qp, sp_, int(k), nr, rp, rsp,
                                     _abi.OTG_MEM_DEVICE if qdev else _abi.OTG_MEM_HOST, C.byref(out))
        _check(st, "otg_autotune", self)
        return {f: getattr(out, f) for f, _ in _abi.Tuning._fields_}

    def set_tuning(self, index_capacity: int = 0, waves_per_simd: int = 0, order_queries: int = 1,
                   persistent: int = 0, range_records: int = 0) -> None:
        t = _abi.Tuning(0, index_capacity, waves_per_simd, order_queries, persistent, range_records, 0.0)
        _check(_abi.lib().otg_ctx_set_tuning(self.handle, C.byref(t)), "otg_ctx_set_tuning", self)

    def tuning(self) -> dict:
        t = _abi.Tuning()
        _check(_abi.lib().otg_ctx_get_tuning(self.handle, C.byref(t)), "otg_ctx_get_tuning", self)
        return {f: getattr(t, f) for f, _ in _abi.Tuning._fields_}

    def mem_stats(self):
        a, b = C.c_uint64(), C.c_uint64()
        _check(_abi.lib().otg_ctx_mem_stats(self.handle, C.byref(a), C.byref(b)), "mem_stats", self)
        return a.value, b.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            for child in list(getattr(self, "_children", ())):
                child.free()
            _abi.lib().otg_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Result:
    """CSR result (otg_result): offsets (nq+1,) uint64, ids uint32."""

    def __init__(self, handle, ctx: Optional[Context]):
        self.handle = handle
        self.ctx = ctx  # None for host-resident results (served requests)
        if ctx is not None:
            ctx._children.add(self)

    def csr(self):
        L = _abi.lib()
        nq, nids = C.c_uint64(), C.c_uint64()
        _check(L.otg_result_get_info(self.handle, C.byref(nq), C.byref(nids)), "result_get_info", self.ctx)
        offsets = np.empty(nq.value + 1, dtype=np.uint64)
        ids = np.empty(max(nids.value, 1), dtype=np.uint32)
        _check(L.otg_result_copy(self.handle, _ptr(offsets), _ptr(ids)), "result_copy", self.ctx)
        return offsets, ids[: nids.value]

    def lists(self):
        off, ids = self.csr()
        return [ids[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)]

    def free(self):
        if self.handle:
            _abi.lib().otg_result_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class KnnPlan:
    """A captured kNN batch (otg_knn_plan): run() replays it on the current contents of
    the query tensors it was made from; csr() / device_csr() read the last run."""

    def __init__(self, handle, tree, nq: int, keep):
        self.handle = handle
        self.tree = tree
        self.nq = nq
        self._keep = keep  # the query tensors the graph reads

    def run(self) -> "KnnPlan":
        _check(_abi.lib().otg_knn_plan_run(self.handle), "otg_knn_plan_run", self.tree.ctx)
        return self

    def device_csr(self):
        """(offsets pointer, ids pointer, id count) of the last run, in device memory."""
        off, ids, n = C.c_void_p(), C.c_void_p(), C.c_uint64()
        _check(_abi.lib().otg_knn_plan_result(self.handle, C.byref(off), C.byref(ids), C.byref(n)),
               "otg_knn_plan_result", self.tree.ctx)
        return off.value, ids.value, n.value

    def csr(self):
        n = self.device_csr()[2]
        offsets = np.empty(self.nq + 1, dtype=np.uint64)
        ids = np.empty(max(n, 1), dtype=np.uint32)
        _check(_abi.lib().otg_knn_plan_copy(self.handle, _ptr(offsets), _ptr(ids)), "otg_knn_plan_copy",
               self.tree.ctx)
        return offsets, ids[:n]

    def free(self):
        if getattr(self, "handle", None):
            _abi.lib().otg_knn_plan_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class RangePlan:
    """A captured RangeSearch batch (otg_range_plan): run() replays it on the current
    contents of the box tensors it was made from; csr() / device_csr() read the last run."""

    def __init__(self, handle, tree, nq: int, keep):
        self.handle = handle
        self.tree = tree
        self.nq = nq
        self._keep = keep

    def run(self) -> "RangePlan":
        _check(_abi.lib().otg_range_plan_run(self.handle), "otg_range_plan_run", self.tree.ctx)
        return self

    def device_csr(self):
        off, ids, n = C.c_void_p(), C.c_void_p(), C.c_uint64()
        _check(_abi.lib().otg_range_plan_result(self.handle, C.byref(off), C.byref(ids), C.byref(n)),
               "otg_range_plan_result", self.tree.ctx)
        return off.value, ids.value, n.value

    def csr(self):
        n = self.device_csr()[2]
        offsets = np.empty(self.nq + 1, dtype=np.uint64)
        ids = np.empty(max(n, 1), dtype=np.uint32)
        _check(_abi.lib().otg_range_plan_copy(self.handle, _ptr(offsets), _ptr(ids)), "otg_range_plan_copy",
               self.tree.ctx)
        return offsets, ids[:n]

    def free(self):
        if getattr(self, "handle", None):
            _abi.lib().otg_range_plan_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BuildPlan:
    """A point-forest build captured as a HIP graph (otg_build_plan): made from device
    points and host set offsets (every set's box given), replayed by run() on whatever
    the points tensor holds then.  `tree` is the plan's tree (queries and kNN plans
    take it; it is freed with the plan)."""

    def __init__(self, points, offsets, max_elements, boxes, max_depths=None, *, ctx=None, dim: int = 3):
        self.ctx = ctx or default_context()
        self.handle = None
        if not _is_device(points) or not points.is_contiguous():
            raise ValueError("a build plan takes a contiguous device tensor of points")
        offs = np.ascontiguousarray(np.asarray(offsets, dtype=np.uint64))
        ns = len(offs) - 1
        caps = np.ascontiguousarray(np.broadcast_to(np.asarray(max_elements, dtype=np.uint32), (ns,)))
        deps = None if max_depths is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(max_depths, dtype=np.uint32), (ns,)))
        bxs = _f64(boxes, 2 * dim)
        h = C.c_void_p()
        st = _abi.lib().otg_build_plan_create(self.ctx.handle, dim, ns, offs.ctypes.data_as(_abi.U64P),
                                              C.c_void_p(points.data_ptr()), _ptr(bxs),
                                              None if deps is None else deps.ctypes.data_as(_abi.U32P),
                                              caps.ctypes.data_as(_abi.U32P), C.byref(h))
        _check(st, "otg_build_plan_create", self.ctx)
        self.handle = h
        self._points = points  # the graph reads this tensor's memory
        th = C.c_void_p()
        _check(_abi.lib().otg_build_plan_tree(h, C.byref(th)), "otg_build_plan_tree", self.ctx)
        t = StaticTree.__new__(StaticTree)
        t.ctx, t.kind, t.dim, t.handle, t.n, t._borrowed = self.ctx, "point", dim, th, int(offs[-1]), True
        self.tree = t
        self._steps = []

    def run(self) -> "BuildPlan":
        _check(_abi.lib().otg_build_plan_run(self.handle), "otg_build_plan_run", self.ctx)
        return self

    def step(self, knn_plan: Optional[KnnPlan] = None, range_plan: Optional[RangePlan] = None) -> "StepPlan":
        """The build, then a kNN plan and / or a range plan made on self.tree, replayed as
        one graph."""
        h = C.c_void_p()
        _check(_abi.lib().otg_step_plan_create(self.handle, knn_plan.handle if knn_plan else None,
                                               range_plan.handle if range_plan else None, C.byref(h)),
               "otg_step_plan_create", self.ctx)
        sp = StepPlan(h, self, knn_plan, range_plan)
        self._steps.append(sp)
        return sp

    def free(self):
        if getattr(self, "handle", None):
            for sp in self._steps:
                sp.free()
            self.tree.free()
            _abi.lib().otg_build_plan_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class StepPlan:
    """A build plan and query plans on its tree as one graph (otg_step_plan); after run()
    the tree is current and each query plan's csr() holds its batch's ids."""

    def __init__(self, handle, build: BuildPlan, knn: Optional[KnnPlan], rng: Optional[RangePlan]):
        self.handle, self.build, self.knn, self.range = handle, build, knn, rng

    def run(self) -> "StepPlan":
        _check(_abi.lib().otg_step_plan_run(self.handle), "otg_step_plan_run", self.build.ctx)
        return self

    def free(self):
        if getattr(self, "handle", None):
            _abi.lib().otg_step_plan_free(self.handle)
            self.handle = None


def _is_device(a) -> bool:
    """Objects exposing data_ptr() (e.g. torch device tensors) are passed as device memory."""
    return hasattr(a, "data_ptr") and getattr(getattr(a, "device", None), "type", "cpu") != "cpu"


def _dev(a, cols: int):
    """(pointer, rows) of a contiguous device float64/int32 tensor."""
    if not a.is_contiguous():
        raise ValueError("device input must be contiguous")
    rows = a.numel() // cols if cols else a.numel()
    return C.c_void_p(a.data_ptr()), rows


def _f64(a, cols: int) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return a.reshape(-1, cols) if a.size else a.reshape(0, cols)


class StaticTree:
    """Static linear ortho-tree resident in HBM (one set) or a forest of sets.

    Mirrors StaticLinearOrthoTreeCore + OrthoTreeQueryBase (core/ot_static_linear_core.h,
    core/ot_query.h).  kind: 'point' or 'box'.
    """

    KIND = "point"
    DIM = 3

    def __init__(self, exec_tag=None, entities=None, max_depth: Optional[int] = None,
                 box_space=None, max_elements: int = 16, *, ctx: Optional[Context] = None,
                 dim: Optional[int] = None, kind: Optional[str] = None):
        self.ctx = ctx or default_context()
        self.kind = kind or self.KIND
        self.dim = dim or self.DIM
        self.handle = None
        if isinstance(exec_tag, str) or exec_tag is None:
            ents = entities
        else:  # constructor without an exec tag: (entities, depth, box, maxElem)
            ents = exec_tag
        if ents is not None:
            ok = self.Create(ents, max_depth, box_space, max_elements)
            assert ok

    # -- build ------------------------------------------------------------
    def _cols(self) -> int:
        return self.dim if self.kind == "point" else 2 * self.dim

    def Create(self, entities, max_depth: Optional[int] = None, box_space=None,
               max_elements: int = 16, exec_tag=PAR_EXEC) -> bool:
        """StaticLinearOrthoTreeCore::Create (ot_static_linear_core.h:523-599): PAR_EXEC
        numbers the tree as Create<ParExec> (sorted locations), SEQ_EXEC as
        Create<SeqExec> (the unsorted partition build, the reference's default)."""
        self.free()
        if _is_device(entities):
            gptr, n = _dev(entities, self._cols())
            mem = _abi.OTG_MEM_DEVICE
        else:
            geom = _f64(entities, self._cols())
            gptr, n, mem = _ptr(geom), geom.shape[0], _abi.OTG_MEM_HOST
        box = None if box_space is None else _f64(box_space, 2 * self.dim).reshape(-1)
        h = C.c_void_p()
        L = _abi.lib()
        kind = _abi.OTG_KIND_POINT if self.kind == "point" else _abi.OTG_KIND_BOX
        fn = L.otg_tree_build_seq if exec_tag == SEQ_EXEC else L.otg_tree_build
        st = fn(self.ctx.handle, kind, self.dim, n, gptr, mem, _ptr(box), int(max_depth or 0), int(max_elements),
                C.byref(h))
        _check(st, "otg_tree_build_seq" if exec_tag == SEQ_EXEC else "otg_tree_build", self.ctx)
        self.handle = h
        self.ctx._children.add(self)
        self.n = n
        return True

    @classmethod
    def forest(cls, sets: Sequence, max_elements, max_depths=None, boxes=None, *, ctx=None, dim=None, kind=None,
               device_geom=None):
        """Batched build of many independent trees (otg_forest_build)."""
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        self.kind = kind or cls.KIND
        self.dim = dim or cls.DIM
        self.handle = None
        cols = self._cols()
        if device_geom is not None:  # (device tensor of all sets, host offsets)
            dgeom, offs = device_geom
            offs = np.ascontiguousarray(np.asarray(offs, dtype=np.uint64))
            ns = len(offs) - 1
            gptr, mem = C.c_void_p(dgeom.data_ptr()), _abi.OTG_MEM_DEVICE
        else:
            arrs = [_f64(s, cols) for s in sets]
            ns = len(arrs)
            offs = np.zeros(ns + 1, dtype=np.uint64)
            offs[1:] = np.cumsum([a.shape[0] for a in arrs])
            geom = np.concatenate(arrs) if arrs else np.zeros((0, cols))
            gptr, mem = _ptr(geom), _abi.OTG_MEM_HOST
        caps = np.ascontiguousarray(np.broadcast_to(np.asarray(max_elements, dtype=np.uint32), (ns,)))
        deps = None if max_depths is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(max_depths, dtype=np.uint32), (ns,)))
        bxs = None if boxes is None else _f64(boxes, 2 * self.dim)
        h = C.c_void_p()
        kindc = _abi.OTG_KIND_POINT if self.kind == "point" else _abi.OTG_KIND_BOX
        st = _abi.lib().otg_forest_build(self.ctx.handle, kindc, self.dim, ns,
                                         offs.ctypes.data_as(_abi.U64P), gptr, mem, _ptr(bxs),
                                         _ptr(deps), caps.ctypes.data_as(_abi.U32P), C.byref(h))
        _check(st, "otg_forest_build", self.ctx)
        self.handle = h
        self.ctx._children.add(self)
        self.n = int(offs[-1])
        return self

    def free(self):
        if getattr(self, "handle", None):
            for plan in list(getattr(self, "_plans", ())):
                plan.free()
            if not getattr(self, "_borrowed", False):  # a build plan's tree is freed with its plan
                _abi.lib().otg_tree_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # -- node API / export ------------------------------------------------
    def info(self) -> _abi.TreeInfo:
        ti = _abi.TreeInfo()
        _check(_abi.lib().otg_tree_get_info(self.handle, C.byref(ti)), "tree_get_info", self.ctx)
        return ti

    def set_info(self, s: int = 0) -> _abi.SetInfo:
        si = _abi.SetInfo()
        _check(_abi.lib().otg_tree_get_set_info(self.handle, s, C.byref(si)), "tree_get_set_info", self.ctx)
        return si

    def GetNodeCount(self, s: int = 0) -> int:
        return int(self.set_info(s).node_count)

    def GetMaxDepthID(self, s: int = 0) -> int:
        return int(self.set_info(s).max_depth)

    def export(self, s: int = 0, sorted_keys: bool = True) -> dict:
        """Node table of set s (otg_tree_export); sorted_keys=False leaves out the
        entities' location keys (a point tree loaded from an archive has none)."""
        si = self.set_info(s)
        nn, ne, D = int(si.node_count), int(si.entity_count), self.dim
        out = {
            "depth": np.empty(nn, np.uint8), "ebegin": np.empty(nn, np.uint32), "ecount": np.empty(nn, np.uint32),
            "cbegin": np.empty(nn, np.uint32), "ccount": np.empty(nn, np.uint32), "key": np.empty(nn, np.uint64),
            "min": np.empty((nn, D), np.float64), "size": np.empty((nn, D), np.float64),
            "ids": np.empty(ne, np.uint32), "sorted_keys": np.empty(ne, np.uint64),
        }
        if not sorted_keys:
            del out["sorted_keys"]
        st = _abi.lib().otg_tree_export(self.handle, s, *[_ptr(out.get(k)) for k in
                                                          ("depth", "ebegin", "ecount", "cbegin", "ccount", "key",
                                                           "min", "size", "ids", "sorted_keys")])
        _check(st, "otg_tree_export", self.ctx)
        out["max_depth"] = int(si.max_depth)
        out["space_min"] = np.array(si.space_min[:D])
        out["space_max"] = np.array(si.space_max[:D])
        return out

    def serialize(self, s: int = 0, large_dataset: bool = False, entities=None) -> bytes:
        """Binary archive of set s in the reference's BinaryOutputArchive layout
        (serialization.h:373-386, serialization/binary_archive.h:85-175).  With
        `entities` (the build input of the set) the archive is the one of the
        owning OrthoTreeManaged tree (serialization.h:433-439): the core
        followed by the entity vector."""
        flags = _abi.OTG_ARCHIVE_LARGE_DATASET if large_dataset else 0
        n = C.c_uint64(0)
        L = _abi.lib()
        _check(L.otg_tree_serialize(self.handle, s, flags, None, 0, C.byref(n)), "otg_tree_serialize", self.ctx)
        buf = np.empty(int(n.value), np.uint8)
        _check(L.otg_tree_serialize(self.handle, s, flags, _ptr(buf), buf.size, C.byref(n)), "otg_tree_serialize",
               self.ctx)
        out = buf.tobytes()
        if entities is not None:
            from .archive import entity_vector_bytes
            out += entity_vector_bytes(_f64(entities, self._cols()), self.dim, self.kind == "box", large_dataset)
        return out

    @classmethod
    def deserialize(cls, archive: bytes, entities=None, large_dataset: bool = False, *, ctx=None, dim=None,
                    kind=None):
        """Tree loaded from a static-core archive (otg_tree_deserialize, the reading side
        of serialization.h:359-386): `entities` in id order (host array or device
        tensor), or None for an OrthoTreeManaged archive that carries them."""
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        self.kind = kind or cls.KIND
        self.dim = dim or cls.DIM
        self.handle = None
        buf = np.frombuffer(bytes(archive), dtype=np.uint8)
        keep = None
        if entities is None:
            gp, n, mem = None, 0, _abi.OTG_MEM_HOST
        elif _is_device(entities):
            gp, n = _dev(entities, self._cols())
            mem, keep = _abi.OTG_MEM_DEVICE, entities
        else:
            keep = _f64(entities, self._cols())
            gp, n, mem = _ptr(keep), keep.shape[0], _abi.OTG_MEM_HOST
        h = C.c_void_p()
        flags = _abi.OTG_ARCHIVE_LARGE_DATASET if large_dataset else 0
        st = _abi.lib().otg_tree_deserialize(self.ctx.handle, _abi.OTG_KIND_POINT if self.kind == "point" else
                                             _abi.OTG_KIND_BOX, self.dim, _ptr(buf), buf.size, flags, n, gp, mem,
                                             C.byref(h))
        _check(st, "otg_tree_deserialize", self.ctx)
        self.handle = h
        self.ctx._children.add(self)
        del keep
        return self

    # -- batched queries (CSR) -------------------------------------------------
    def _sets_arg(self, set_ids, nq):
        if set_ids is None:
            return None
        if _is_device(set_ids):
            return set_ids
        return np.ascontiguousarray(np.broadcast_to(np.asarray(set_ids, dtype=np.uint32), (nq,)))

    def _query_args(self, q, cols, set_ids):
        """(pointer, nq, set pointer, mem, keepalive) for host arrays or device tensors."""
        if _is_device(q):
            qp, nq = _dev(q, cols)
            sp = None if set_ids is None else C.c_void_p(set_ids.data_ptr())
            return qp, nq, sp, _abi.OTG_MEM_DEVICE, (q, set_ids)
        qa = _f64(q, cols)
        sids = self._sets_arg(set_ids, qa.shape[0])
        return _ptr(qa), qa.shape[0], _ptr(sids), _abi.OTG_MEM_HOST, (qa, sids)

    def knn_batch(self, points, k: int, max_within: float = DBL_MAX, tolerance: float = BASE_TOLERANCE,
                  set_ids=None) -> Result:
        qp, nq, sp, mem, _keep = self._query_args(points, self.dim, set_ids)
        h = C.c_void_p()
        st = _abi.lib().otg_knn(self.handle, nq, qp, sp, mem, int(k), float(max_within), float(tolerance), C.byref(h))
        _check(st, "otg_knn", self.ctx)
        return Result(h, self.ctx)

    def knn_plan(self, points, k: int, max_within: float = DBL_MAX, tolerance: float = BASE_TOLERANCE,
                 set_ids=None) -> "KnnPlan":
        """otg_knn_plan_create: the kNN batch over these device tensors, captured once
        and replayed by KnnPlan.run() on whatever they hold at that time."""
        if not _is_device(points):
            raise ValueError("a kNN plan takes device-resident query points")
        qp, nq = _dev(points, self.dim)
        if set_ids is not None and (not _is_device(set_ids) or not set_ids.is_contiguous()):
            raise ValueError("set ids of a plan must be a contiguous device tensor")
        sp = None if set_ids is None else C.c_void_p(set_ids.data_ptr())
        h = C.c_void_p()
        st = _abi.lib().otg_knn_plan_create(self.handle, nq, qp, sp, int(k), float(max_within), float(tolerance),
                                            C.byref(h))
        _check(st, "otg_knn_plan_create", self.ctx)
        plan = KnnPlan(h, self, nq, (points, set_ids))
        if not hasattr(self, "_plans"):
            self._plans = weakref.WeakSet()
        self._plans.add(plan)
        return plan

    def range_plan(self, ranges, mode: int = _abi.OTG_RANGE_INSIDE, tolerance: float = BASE_TOLERANCE,
                   set_ids=None) -> "RangePlan":
        """otg_range_plan_create: the RangeSearch batch over these device tensors (boxes as
        min then max, 2 x dim columns), captured once and replayed by RangePlan.run()."""
        if not _is_device(ranges):
            raise ValueError("a range plan takes device-resident boxes")
        rp, nq = _dev(ranges, 2 * self.dim)
        if set_ids is not None and (not _is_device(set_ids) or not set_ids.is_contiguous()):
            raise ValueError("set ids of a plan must be a contiguous device tensor")
        sp = None if set_ids is None else C.c_void_p(set_ids.data_ptr())
        h = C.c_void_p()
        st = _abi.lib().otg_range_plan_create(self.handle, nq, rp, sp, int(mode), float(tolerance), C.byref(h))
        _check(st, "otg_range_plan_create", self.ctx)
        plan = RangePlan(h, self, nq, (ranges, set_ids))
        if not hasattr(self, "_plans"):
            self._plans = weakref.WeakSet()
        self._plans.add(plan)
        return plan

    def knn_footprint(self, points, k: int, max_within: float = DBL_MAX, tolerance: float = BASE_TOLERANCE,
                      set_ids=None) -> dict:
        """Minimal set the kNN batch must touch (otg_knn_footprint): node visits with wall
        distance below each query's final radius, the entities they hold, and result ids."""
        qp, nq, sp, mem, _keep = self._query_args(points, self.dim, set_ids)
        out = (C.c_uint64 * 3)()
        st = _abi.lib().otg_knn_footprint(self.handle, nq, qp, sp, mem, int(k), float(max_within), float(tolerance),
                                          out)
        _check(st, "otg_knn_footprint", self.ctx)
        return {"nodes": int(out[0]), "entities": int(out[1]), "ids": int(out[2])}

    def range_batch(self, ranges, mode: int = _abi.OTG_RANGE_INSIDE, tolerance: float = BASE_TOLERANCE,
                    set_ids=None) -> Result:
        qp, nq, sp, mem, _keep = self._query_args(ranges, 2 * self.dim, set_ids)
        h = C.c_void_p()
        st = _abi.lib().otg_range(self.handle, nq, qp, sp, mem, int(mode), float(tolerance), C.byref(h))
        _check(st, "otg_range", self.ctx)
        return Result(h, self.ctx)

    def ray_first_batch(self, rays, tolerance: float = BASE_TOLERANCE, tolerance_increment: float = 0.0,
                        max_distance: float = DBL_MAX, set_ids=None) -> Result:
        qp, nq, sp, mem, _keep = self._query_args(rays, 2 * self.dim, set_ids)
        h = C.c_void_p()
        st = _abi.lib().otg_ray_first(self.handle, nq, qp, sp, mem, float(tolerance), float(tolerance_increment),
                                      float(max_distance), C.byref(h))
        _check(st, "otg_ray_first", self.ctx)
        return Result(h, self.ctx)

    def pick_batch(self, points, tolerance: float = BASE_TOLERANCE, set_ids=None) -> Result:
        qp, nq, sp, mem, _keep = self._query_args(points, self.dim, set_ids)
        h = C.c_void_p()
        st = _abi.lib().otg_pick(self.handle, nq, qp, sp, mem, float(tolerance), C.byref(h))
        _check(st, "otg_pick", self.ctx)
        return Result(h, self.ctx)

    def planes_batch(self, planes, mode: int = _abi.OTG_PLANE_SEARCH, tolerance: float = BASE_TOLERANCE,
                     set_ids=None) -> Result:
        """planes: (nq, P, 1 + D) -- origo distance, unit normal -- or (nq, 1 + D) for one
        plane per query.  mode: OTG_PLANE_SEARCH, OTG_PLANE_POSITIVE or OTG_PLANE_FRUSTUM."""
        if _is_device(planes):
            nplanes = 1 if planes.dim() == 2 else planes.shape[1]
            qp, nq = _dev(planes.reshape(planes.shape[0], -1), nplanes * (self.dim + 1))
            sp = None if set_ids is None else C.c_void_p(set_ids.data_ptr())
            mem, keep = _abi.OTG_MEM_DEVICE, (planes, set_ids)
        else:
            a = np.asarray(planes, dtype=np.float64)
            if a.ndim == 2:
                a = a[:, None, :]
            nplanes = a.shape[1]
            qa = np.ascontiguousarray(a.reshape(a.shape[0], -1))
            nq = qa.shape[0]
            sids = self._sets_arg(set_ids, nq)
            qp, sp, mem, keep = _ptr(qa), _ptr(sids), _abi.OTG_MEM_HOST, (qa, sids)
        h = C.c_void_p()
        st = _abi.lib().otg_planes(self.handle, nq, int(nplanes), qp, sp, mem, int(mode), float(tolerance), C.byref(h))
        _check(st, "otg_planes", self.ctx)
        return Result(h, self.ctx)

    def ray_all_batch(self, rays, tolerance: float = BASE_TOLERANCE, tolerance_increment: float = 0.0,
                      max_distance: float = DBL_MAX, set_ids=None) -> Result:
        qp, nq, sp, mem, _keep = self._query_args(rays, 2 * self.dim, set_ids)
        h = C.c_void_p()
        st = _abi.lib().otg_ray_all(self.handle, nq, qp, sp, mem, float(tolerance), float(tolerance_increment),
                                    float(max_distance), C.byref(h))
        _check(st, "otg_ray_all", self.ctx)
        return Result(h, self.ctx)

    def collide_with(self, other: "StaticTree", tolerance: float = BASE_TOLERANCE, s: int = 0,
                     other_set: int = 0) -> np.ndarray:
        """Two-tree collision: (id in this tree, id in other) pairs, sorted."""
        h = C.c_void_p()
        _check(_abi.lib().otg_collide_trees(self.handle, s, other.handle, other_set, float(tolerance), C.byref(h)),
               "otg_collide_trees", self.ctx)
        r = Result(h, self.ctx)
        _, ids = r.csr()
        r.free()
        return ids.reshape(-1, 2)

    def collide(self, tolerance: float = BASE_TOLERANCE, s: int = 0) -> np.ndarray:
        h = C.c_void_p()
        _check(_abi.lib().otg_collide(self.handle, s, float(tolerance), C.byref(h)), "otg_collide", self.ctx)
        r = Result(h, self.ctx)
        _, ids = r.csr()
        r.free()
        return ids.reshape(-1, 2)

    # -- reference-named single queries (ot_query.h) ----------------------
    def GetNearestNeighbors(self, point, k: int, *args, tolerance: float = BASE_TOLERANCE):
        """GetNearestNeighbors(point, k, entities, tol) or (point, k, maxDistanceWithin, entities, tol)."""
        max_within = DBL_MAX
        rest = list(args)
        if rest and np.isscalar(rest[0]):
            max_within = float(rest.pop(0))
        if rest:
            rest.pop(0)  # entities span (resident on the device already)
        if rest:
            tolerance = float(rest.pop(0))
        return self.knn_batch([point], k, max_within, tolerance).lists()[0]

    def RangeSearch(self, box, entities=None, mode: int = _abi.OTG_RANGE_INSIDE, tolerance: float = BASE_TOLERANCE):
        return self.range_batch([np.asarray(box, dtype=np.float64).reshape(-1)], mode, tolerance).lists()[0]

    def RayIntersectedFirst(self, origin, direction, entities=None, tolerance: float = BASE_TOLERANCE,
                            tolerance_increment: float = 0.0, max_distance: float = DBL_MAX):
        ray = np.concatenate([np.asarray(origin, np.float64).reshape(-1), np.asarray(direction, np.float64).reshape(-1)])
        return self.ray_first_batch([ray], tolerance, tolerance_increment, max_distance).lists()[0]

    def PickSearch(self, pick_point, entities=None, tolerance: float = BASE_TOLERANCE):
        return self.pick_batch([np.asarray(pick_point, dtype=np.float64).reshape(-1)], tolerance).lists()[0]

    def PlaneSearch(self, distance_of_origo: float, plane_normal, entities=None, tolerance: float = BASE_TOLERANCE):
        pl = np.concatenate([[float(distance_of_origo)], np.asarray(plane_normal, np.float64).reshape(-1)])
        return self.planes_batch([pl], _abi.OTG_PLANE_SEARCH, tolerance).lists()[0]

    PlaneIntersection = PlaneSearch

    def PlanePositiveSegmentation(self, distance_of_origo: float, plane_normal, entities=None,
                                  tolerance: float = BASE_TOLERANCE):
        pl = np.concatenate([[float(distance_of_origo)], np.asarray(plane_normal, np.float64).reshape(-1)])
        return self.planes_batch([pl], _abi.OTG_PLANE_POSITIVE, tolerance).lists()[0]

    def FrustumCulling(self, boundary_planes, entities=None, tolerance: float = BASE_TOLERANCE):
        """boundary_planes: (P, 1 + D) rows of (origo distance, unit normal)."""
        pl = np.asarray(boundary_planes, np.float64).reshape(1, -1, self.dim + 1)
        if pl.shape[1] == 0:
            return []
        return self.planes_batch(pl, _abi.OTG_PLANE_FRUSTUM, tolerance).lists()[0]

    def RayIntersectedAll(self, origin, direction, entities=None, tolerance: float = BASE_TOLERANCE,
                          tolerance_increment: float = 0.0, max_distance: float = DBL_MAX):
        ray = np.concatenate([np.asarray(origin, np.float64).reshape(-1), np.asarray(direction, np.float64).reshape(-1)])
        return self.ray_all_batch([ray], tolerance, tolerance_increment, max_distance).lists()[0]

    def CollisionDetection(self, *args, tolerance: float = BASE_TOLERANCE, exec_tag=PAR_EXEC):
        """CollisionDetection(entities, tol) on this tree (ot_query.h:1901-1906), or
        CollisionDetection(entities, otherTree, otherEntities, tol) against another tree
        (:1488-1496): pairs (this id, other id)."""
        rest = list(args)
        if len(rest) >= 2 and isinstance(rest[1], StaticTree):
            other = rest[1]
            if len(rest) >= 4:
                tolerance = float(rest[3])
            return [tuple(p) for p in self.collide_with(other, tolerance).tolist()]
        if len(rest) >= 2:
            tolerance = float(rest[1])
        return [tuple(p) for p in self.collide(tolerance).tolist()]

    @staticmethod
    def CollisionDetectionTrees(left: "StaticTree", left_entities, right: "StaticTree", right_entities,
                                tolerance: float = BASE_TOLERANCE):
        """Static CollisionDetection(leftTree, leftEntities, rightTree, rightEntities, tol)
        (ot_query.h:1462-1484)."""
        return [tuple(p) for p in left.collide_with(right, tolerance).tolist()]


class Server:
    """Resident trees under an HBM budget with continuous batching (otg_server_*).

    Requests are submitted without blocking and collected with wait(); a served
    request's ids equal the direct StaticTree query's on the same queries.
    """

    def __init__(self, device: int = 0, hbm_budget: int = 64 << 30, workers: int = 2,
                 batch_queries: Optional[int] = None):
        h = C.c_void_p()
        st = _abi.lib().otg_server_create(device, int(hbm_budget), int(workers), C.byref(h))
        if st != _abi.OTG_OK:
            raise OtgError(st, "otg_server_create")
        self.handle = h
        self._dims = {}
        if batch_queries is not None:
            self.set_batch_limit(batch_queries)

    def set_batch_limit(self, max_queries: int) -> None:
        """The most queries one batch takes (otg_server_set_batch_limit): smaller batches
        lower the tail latency at a given load and cost throughput."""
        _check(_abi.lib().otg_server_set_batch_limit(self.handle, int(max_queries)), "otg_server_set_batch_limit")

    def add_tree(self, tree_id: int, kind: str, dim: int, entities, max_depth: Optional[int] = None,
                 box_space=None, max_elements: int = 16) -> None:
        cols = dim if kind == "point" else 2 * dim
        geom = _f64(entities, cols)
        box = None if box_space is None else _f64(box_space, 2 * dim).reshape(-1)
        kindc = _abi.OTG_KIND_POINT if kind == "point" else _abi.OTG_KIND_BOX
        st = _abi.lib().otg_server_add_tree(self.handle, int(tree_id), kindc, dim, geom.shape[0], _ptr(geom),
                                            _abi.OTG_MEM_HOST, _ptr(box), int(max_depth or 0), int(max_elements))
        _check(st, "otg_server_add_tree")
        self._dims[int(tree_id)] = (kind, dim)

    def remove_tree(self, tree_id: int) -> None:
        _check(_abi.lib().otg_server_remove_tree(self.handle, int(tree_id)), "otg_server_remove_tree")
        self._dims.pop(int(tree_id), None)

    def tree_state(self, tree_id: int) -> int:
        """0 not registered, 1 resident, 2 registered but evicted (rebuilt on demand)."""
        v = C.c_int()
        _check(_abi.lib().otg_server_has_tree(self.handle, int(tree_id), C.byref(v)), "otg_server_has_tree")
        return v.value

    def has_tree(self, tree_id: int) -> bool:
        """True when the tree is resident in HBM."""
        return self.tree_state(tree_id) == 1

    def _dim(self, tree_id) -> int:
        if int(tree_id) not in self._dims:
            raise OtgError(1, "otg_server_submit", f"tree {tree_id} is not registered")
        return self._dims[int(tree_id)][1]

    def _submit(self, tree_id, query, q, k=0, p0=0.0, p1=0.0, p2=0.0, mode=0) -> int:
        rid = C.c_uint64()
        st = _abi.lib().otg_server_submit(self.handle, int(tree_id), query, q.shape[0], _ptr(q), int(k), float(p0),
                                          float(p1), float(p2), int(mode), C.byref(rid))
        _check(st, "otg_server_submit")
        return rid.value

    def submit_knn(self, tree_id: int, points, k: int, max_within: float = DBL_MAX,
                   tolerance: float = BASE_TOLERANCE) -> int:
        return self._submit(tree_id, 0, _f64(points, self._dim(tree_id)), k=k, p0=max_within, p1=tolerance)

    def submit_range(self, tree_id: int, ranges, mode: int = _abi.OTG_RANGE_INSIDE,
                     tolerance: float = BASE_TOLERANCE) -> int:
        return self._submit(tree_id, 1, _f64(ranges, 2 * self._dim(tree_id)), p0=tolerance, mode=mode)

    def submit_ray(self, tree_id: int, rays, tolerance: float = BASE_TOLERANCE, tolerance_increment: float = 0.0,
                   max_distance: float = DBL_MAX) -> int:
        return self._submit(tree_id, 2, _f64(rays, 2 * self._dim(tree_id)), p0=tolerance,
                            p1=tolerance_increment, p2=max_distance)

    def wait(self, request_id: int):
        """-> (Result, latency_ms, queue_ms)."""
        h, lat, qms = C.c_void_p(), C.c_double(), C.c_double()
        st = _abi.lib().otg_server_wait(self.handle, int(request_id), C.byref(h), C.byref(lat), C.byref(qms))
        if st != _abi.OTG_OK and h.value:
            _abi.lib().otg_result_free(h)
        _check(st, "otg_server_wait")
        return Result(h, None), lat.value, qms.value

    def wait_detail(self, request_id: int):
        """-> (Result, dict): otg_server_wait_detail -- latency, queueing, service, time
        making trees resident in the batch, batch size, and whether the tree was cold."""
        h, d = C.c_void_p(), (C.c_double * 6)()
        st = _abi.lib().otg_server_wait_detail(self.handle, int(request_id), C.byref(h), d)
        if st != _abi.OTG_OK and h.value:
            _abi.lib().otg_result_free(h)
        _check(st, "otg_server_wait_detail")
        keys = ("latency_ms", "queue_ms", "service_ms", "resident_ms", "batch_requests", "cold")
        return Result(h, None), {k: float(d[i]) for i, k in enumerate(keys)}

    def stats(self) -> dict:
        v = (C.c_uint64 * 15)()
        _check(_abi.lib().otg_server_stats(self.handle, v), "otg_server_stats")
        keys = ["batches", "launches", "requests", "queries", "evictions", "resident_bytes", "resident_trees",
                "pending", "rebuilds", "registered_trees", "build_us", "query_us", "stage_us", "launch_us",
                "split_us"]
        return dict(zip(keys, [int(x) for x in v]))

    def close(self) -> None:
        if getattr(self, "handle", None):
            _abi.lib().otg_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _make(kind: str, dim: int, name: str):
    return type(name, (StaticTree,), {"KIND": kind, "DIM": dim})


# core/ot_aliases.h:225-245
StaticQuadtreePoint = _make("point", 2, "StaticQuadtreePoint")
StaticQuadtreeBox = _make("box", 2, "StaticQuadtreeBox")
StaticOctreePoint = _make("point", 3, "StaticOctreePoint")
StaticOctreeBox = _make("box", 3, "StaticOctreeBox")
StaticHexatreePoint = _make("point", 4, "StaticHexatreePoint")
StaticHexatreeBox = _make("box", 4, "StaticHexatreeBox")


def StaticOrthoTreePointND(dim: int):
    """core/ot_aliases.h:82-85"""
    return _make("point", dim, f"StaticOrthoTreePointND_{dim}")


def StaticOrthoTreeBoxND(dim: int, loose: bool = True):
    """core/ot_aliases.h:88-91 (loose trees only: LOOSE_FACTOR 2.0)"""
    if not loose:
        raise NotImplementedError("non-loose box trees are outside the GPU path")
    return _make("box", dim, f"StaticOrthoTreeBoxND_{dim}")


__all__ = [
    "Context", "Result", "KnnPlan", "RangePlan", "BuildPlan", "StepPlan", "StaticTree", "Server", "OtgError", "BASE_TOLERANCE", "DBL_MAX", "PAR_EXEC", "SEQ_EXEC", "GPU_EXEC",
    "StaticQuadtreePoint", "StaticQuadtreeBox", "StaticOctreePoint", "StaticOctreeBox", "StaticHexatreePoint",
    "StaticHexatreeBox", "StaticOrthoTreePointND", "StaticOrthoTreeBoxND", "default_context",
]
