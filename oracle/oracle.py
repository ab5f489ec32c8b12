"""ctypes binding of the CPU oracle (oracle/liboc.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module, as the checker.  The product never calls it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboc.so"
DBL_MAX = float(np.finfo(np.float64).max)
BASE_TOLERANCE = float(np.finfo(np.float64).eps * 100.0)

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    L = C.CDLL(str(LIB_PATH))
    P = C.c_void_p
    L.oc_build.restype = P
    L.oc_build.argtypes = [C.c_int, C.c_int, C.c_uint32, P, P, C.c_uint32, C.c_uint32]
    L.oc_build_seq.restype = P
    L.oc_build_seq.argtypes = [C.c_int, C.c_int, C.c_uint32, P, P, C.c_uint32, C.c_uint32]
    L.oc_partition_codes.restype = C.c_uint32
    L.oc_partition_codes.argtypes = [P, P, C.c_uint32, C.c_uint64, C.c_uint64]
    L.oc_build_mt.restype = P
    L.oc_build_mt.argtypes = [C.c_int, C.c_int, C.c_uint32, P, P, C.c_uint32, C.c_uint32, C.c_int]
    L.oc_tree_free.argtypes = [P]
    L.oc_tree_status.restype = C.c_int
    L.oc_tree_status.argtypes = [P]
    L.oc_tree_node_count.restype = C.c_uint32
    L.oc_tree_node_count.argtypes = [P]
    L.oc_tree_max_depth.restype = C.c_uint32
    L.oc_tree_max_depth.argtypes = [P]
    L.oc_tree_export.argtypes = [P] + [P] * 12
    for nm in ("oc_knn",):
        getattr(L, nm).restype = P
    L.oc_knn.argtypes = [P, P, C.c_uint64, P, C.c_uint32, C.c_double, C.c_double, C.c_int]
    L.oc_range.restype = P
    L.oc_range.argtypes = [P, P, C.c_uint64, P, C.c_int, C.c_double, C.c_int]
    L.oc_ray_first.restype = P
    L.oc_ray_first.argtypes = [P, P, C.c_uint64, P, C.c_double, C.c_double, C.c_double, C.c_int]
    L.oc_collision.restype = P
    L.oc_collision.argtypes = [P, P, C.c_double]
    L.oc_collision_trees.restype = P
    L.oc_collision_trees.argtypes = [P, P, P, P, C.c_double]
    L.oc_pick.restype = P
    L.oc_pick.argtypes = [P, P, C.c_uint64, P, C.c_double, C.c_int]
    L.oc_planes.restype = P
    L.oc_planes.argtypes = [P, P, C.c_uint64, C.c_uint32, P, C.c_int, C.c_double, C.c_int]
    L.oc_ray_all.restype = P
    L.oc_ray_all.argtypes = [P, P, C.c_uint64, P, C.c_double, C.c_double, C.c_double, C.c_int]
    L.oc_result_nids.restype = C.c_uint64
    L.oc_result_nids.argtypes = [P]
    L.oc_result_nq.restype = C.c_uint64
    L.oc_result_nq.argtypes = [P]
    L.oc_result_copy.argtypes = [P, P, P]
    L.oc_result_free.argtypes = [P]
    L.oc_morton_encode.restype = C.c_uint64
    L.oc_morton_encode.argtypes = [P, C.c_int]
    L.oc_morton_node_key.restype = C.c_uint64
    L.oc_morton_node_key.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int]
    L.oc_morton_decode.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, P]
    L.oc_estimate_max_depth.restype = C.c_uint32
    L.oc_estimate_max_depth.argtypes = [C.c_uint64, C.c_uint64, C.c_int]
    L.oc_bounding_box_export.argtypes = [P, C.c_uint32, C.c_int, C.c_int, P, P]
    L.oc_splitmix_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double, P]
    _lib = L
    return L


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _result(h):
    L = lib()
    nq = L.oc_result_nq(h)
    nids = L.oc_result_nids(h)
    off = np.empty(nq + 1, np.uint64)
    ids = np.empty(max(nids, 1), np.uint32)
    L.oc_result_copy(h, _p(off), _p(ids))
    L.oc_result_free(h)
    return off, ids[:nids]


def morton_encode(g, D):
    a = np.ascontiguousarray(np.asarray(g, dtype=np.uint32))
    return int(lib().oc_morton_encode(_p(a), D))


def node_key(code, depth, max_depth, D):
    return int(lib().oc_morton_node_key(code, depth, max_depth, D))


def decode(key, depth, max_depth, D):
    g = np.zeros(D, np.uint32)
    lib().oc_morton_decode(key, depth, max_depth, D, _p(g))
    return g.tolist()


def estimate_max_depth(n, cap, D):
    return int(lib().oc_estimate_max_depth(n, cap, D))


class OracleTree:
    """CPU restatement of StaticLinearOrthoTreeCore (points: MinPoint, boxes: loose 2.0 + MBR)."""

    def __init__(self, dim: int, kind: str, geom, box=None, max_depth: int = 0, cap: int = 16, threads: int = 1,
                 seq: bool = False):
        self.dim = dim
        self.kind = kind
        cols = dim if kind == "point" else 2 * dim
        self.geom = np.ascontiguousarray(np.asarray(geom, np.float64).reshape(-1, cols))
        bx = None if box is None else np.ascontiguousarray(np.asarray(box, np.float64).reshape(-1))
        if seq:  # Create<SeqExec>: the unsorted partition build
            self.h = lib().oc_build_seq(dim, 0 if kind == "point" else 1, self.geom.shape[0], _p(self.geom), _p(bx),
                                        int(max_depth or 0), int(cap))
        else:
            self.h = lib().oc_build_mt(dim, 0 if kind == "point" else 1, self.geom.shape[0], _p(self.geom), _p(bx),
                                       int(max_depth or 0), int(cap), int(threads))
        if not self.h:
            raise ValueError("oracle rejected the build parameters")
        self.status = lib().oc_tree_status(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_tree_free(self.h)
            self.h = None

    @property
    def max_depth(self):
        return int(lib().oc_tree_max_depth(self.h))

    def export(self) -> dict:
        nn = int(lib().oc_tree_node_count(self.h))
        n, D = self.geom.shape[0], self.dim
        out = {
            "depth": np.empty(nn, np.uint8), "ebegin": np.empty(nn, np.uint32), "ecount": np.empty(nn, np.uint32),
            "cbegin": np.empty(nn, np.uint32), "ccount": np.empty(nn, np.uint32), "key": np.empty(nn, np.uint64),
            "min": np.empty((nn, D), np.float64), "size": np.empty((nn, D), np.float64),
            "ids": np.empty(n, np.uint32), "sorted_code": np.empty(n, np.uint64), "sorted_depth": np.empty(n, np.uint32),
            "space": np.empty(2 * D, np.float64),
        }
        lib().oc_tree_export(self.h, *[_p(out[k]) for k in ("depth", "ebegin", "ecount", "cbegin", "ccount", "key",
                                                            "min", "size", "ids", "sorted_code", "sorted_depth",
                                                            "space")])
        out["max_depth"] = self.max_depth
        return out

    def knn(self, queries, k, max_within=DBL_MAX, tol=BASE_TOLERANCE, threads=1):
        q = np.ascontiguousarray(np.asarray(queries, np.float64).reshape(-1, self.dim))
        return _result(lib().oc_knn(self.h, _p(self.geom), q.shape[0], _p(q), int(k), float(max_within), float(tol),
                                    int(threads)))

    def range(self, ranges, overlap=False, tol=BASE_TOLERANCE, threads=1):
        q = np.ascontiguousarray(np.asarray(ranges, np.float64).reshape(-1, 2 * self.dim))
        return _result(lib().oc_range(self.h, _p(self.geom), q.shape[0], _p(q), 1 if overlap else 0, float(tol),
                                      int(threads)))

    def ray_first(self, rays, tol=BASE_TOLERANCE, inc=0.0, max_dist=DBL_MAX, threads=1):
        q = np.ascontiguousarray(np.asarray(rays, np.float64).reshape(-1, 2 * self.dim))
        return _result(lib().oc_ray_first(self.h, _p(self.geom), q.shape[0], _p(q), float(tol), float(inc),
                                          float(max_dist), int(threads)))

    def collide(self, tol=BASE_TOLERANCE):
        off, ids = _result(lib().oc_collision(self.h, _p(self.geom), float(tol)))
        return ids.reshape(-1, 2)

    def collide_with(self, other: "OracleTree", tol=BASE_TOLERANCE):
        """Two-tree CollisionDetection: (self id, other id) pairs, sorted."""
        off, ids = _result(lib().oc_collision_trees(self.h, _p(self.geom), other.h, _p(other.geom), float(tol)))
        return ids.reshape(-1, 2)

    def pick(self, points, tol=BASE_TOLERANCE, threads=1):
        q = np.ascontiguousarray(np.asarray(points, np.float64).reshape(-1, self.dim))
        return _result(lib().oc_pick(self.h, _p(self.geom), q.shape[0], _p(q), float(tol), int(threads)))

    def planes(self, planes, mode, tol=BASE_TOLERANCE, threads=1):
        """planes: (nq, P, 1 + D) = origo distance, normal; mode 0 PlaneSearch,
        1 PlanePositiveSegmentation (P = 1), 2 FrustumCulling."""
        q = np.ascontiguousarray(np.asarray(planes, np.float64))
        if q.ndim == 2:
            q = q[:, None, :]
        return _result(lib().oc_planes(self.h, _p(self.geom), q.shape[0], q.shape[1], _p(q), int(mode), float(tol),
                                       int(threads)))

    def ray_all(self, rays, tol=BASE_TOLERANCE, inc=0.0, max_dist=DBL_MAX, threads=1):
        q = np.ascontiguousarray(np.asarray(rays, np.float64).reshape(-1, 2 * self.dim))
        return _result(lib().oc_ray_all(self.h, _p(self.geom), q.shape[0], _p(q), float(tol), float(inc),
                                        float(max_dist), int(threads)))


def lists(off, ids):
    return [ids[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)]


def partition_codes(codes, ids, mask, flag):
    """Partitioning::partition (detail/partitioning.h:215-241) of (code, id) pairs by
    (code & mask) == flag; returns (codes, ids, front count)."""
    c = np.ascontiguousarray(np.asarray(codes, np.uint64)).copy()
    i = np.ascontiguousarray(np.asarray(ids, np.uint32)).copy()
    m = lib().oc_partition_codes(_p(c), _p(i), len(c), int(mask), int(flag))
    return c, i, int(m)
