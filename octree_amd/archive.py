"""Reader and entity writer for the reference's binary archive of a static tree.

The reference serializes a StaticLinearOrthoTreeCore through its
BinaryOutputArchive (serialization/binary_archive.h:85-175): arithmetic values
little-endian in their own width, and every std::vector, std::array and
std::variant index preceded by a count of type serialized_size_t (uint32, or
uint64 when ORTHOTREE__LARGE_DATASET is defined, core/types.h:51-55).  The core
writes (serialization.h:373-386, base :359-371):

    u32 core version, u32 base version, u64 m_maxElementNum, u32 m_maxDepthID,
    box trees:   m_nominalTreeBox = array minPoint, array size
    point trees: m_nodeSize = vector of array (max depth + 2 entries)
    m_nodes = variant index + {vector nodeChildSegmentBegins, vector Segment{begin, length}}
              with NodeStorage256 / 65536 / General widths u8 / u16 / u32 (core/types.h:150-178)
    m_nodeDepthIDs (vector u8), m_entityStorage (vector u32),
    m_nodeGeometry = vector of array minPoint (points) or {minPoint, size} (boxes)

An OrthoTreeManaged archive (serialization.h:433-439) appends the entity
vector: points as arrays, boxes as BoundingBoxND {Min, Max} (:49-55).
"""
from __future__ import annotations

import struct

import numpy as np

_INDEX = {0: "<B", 1: "<H", 2: "<I"}
_SEG = {0: "<B", 1: "<H", 2: "<I"}


class ArchiveError(ValueError):
    pass


class _Reader:
    def __init__(self, data: bytes, large: bool):
        self.b = memoryview(data)
        self.p = 0
        self.large = large

    def take(self, fmt: str):
        size = struct.calcsize(fmt)
        if self.p + size > len(self.b):
            raise ArchiveError(f"archive ends at byte {len(self.b)}, {size} more needed at {self.p}")
        v = struct.unpack_from(fmt, self.b, self.p)
        self.p += size
        return v[0] if len(v) == 1 else v

    def count(self) -> int:
        return int(self.take("<Q" if self.large else "<I"))

    def array(self, dim: int) -> np.ndarray:
        n = self.count()
        if n != dim:
            raise ArchiveError(f"array of {n} coordinates where {dim} were expected at byte {self.p}")
        return np.array(self.take(f"<{dim}d"), dtype=np.float64).reshape(dim)

    def block(self, n: int, fmt: str) -> np.ndarray:
        dt = np.dtype(fmt)
        size = n * dt.itemsize
        if self.p + size > len(self.b):
            raise ArchiveError(f"archive ends inside a block of {n} values at byte {self.p}")
        a = np.frombuffer(self.b, dtype=dt, count=n, offset=self.p).copy()
        self.p += size
        return a


def read_static_archive(data: bytes, dim: int, is_box: bool, large_dataset: bool = False,
                        managed: bool = False) -> dict:
    """Fields of a static-core archive (and the entities of a managed one).

    Returns a dict with max_elements, max_depth, tree_box (box trees),
    node_size (point trees; max depth + 2 rows), storage (variant index),
    child_begins, entity_begin, entity_count, depth, entity_ids, node_min and
    node_size_mbr (box trees), entities (managed archives)."""
    r = _Reader(data, large_dataset)
    out = {}
    if r.take("<I") != 0 or r.take("<I") != 0:
        raise ArchiveError("unknown SERIALIZED_VERSION_ID")
    out["max_elements"] = int(r.take("<Q"))
    out["max_depth"] = int(r.take("<I"))
    if is_box:
        out["tree_box"] = (r.array(dim), r.array(dim))
    else:
        ns = r.count()
        out["node_size"] = np.stack([r.array(dim) for _ in range(ns)]) if ns else np.zeros((0, dim))
    kind = r.count()
    if kind not in _INDEX:
        raise ArchiveError(f"node storage variant index {kind}")
    out["storage"] = kind
    nb = r.count()
    out["child_begins"] = r.block(nb, _INDEX[kind]).astype(np.uint64)
    nseg = r.count()
    seg = "<Q" if (kind == 2 and large_dataset) else _SEG[kind]
    s = r.block(2 * nseg, seg).astype(np.uint64).reshape(nseg, 2)
    out["entity_begin"], out["entity_count"] = s[:, 0], s[:, 1]
    nd = r.count()
    out["depth"] = r.block(nd, "<u1")
    ne = r.count()
    out["entity_ids"] = r.block(ne, "<u4")
    ng = r.count()
    mins, sizes = np.zeros((ng, dim)), np.zeros((ng, dim))
    for i in range(ng):
        mins[i] = r.array(dim)
        if is_box:
            sizes[i] = r.array(dim)
    out["node_min"] = mins
    if is_box:
        out["node_size_mbr"] = sizes
    if not (nseg == nd == ng):
        raise ArchiveError(f"node vectors disagree: {nseg} segments, {nd} depths, {ng} geometries")
    if managed:
        n = r.count()
        if is_box:
            out["entities"] = np.stack([np.concatenate([r.array(dim), r.array(dim)]) for _ in range(n)]) \
                if n else np.zeros((0, 2 * dim))
        else:
            out["entities"] = np.stack([r.array(dim) for _ in range(n)]) if n else np.zeros((0, dim))
    if r.p != len(r.b):
        raise ArchiveError(f"{len(r.b) - r.p} bytes after the archive")
    return out


def entity_vector_bytes(geom: np.ndarray, dim: int, is_box: bool, large_dataset: bool = False) -> bytes:
    """std::vector<Entity> of a managed tree (stl/vector.h; points as
    std::array<double, D>, boxes as BoundingBoxND {Min, Max}, serialization.h:49-55)."""
    geom = np.ascontiguousarray(np.asarray(geom, dtype="<f8").reshape(-1, 2 * dim if is_box else dim))
    n = geom.shape[0]
    cnt = "<Q" if large_dataset else "<I"
    if not large_dataset and n > 0xFFFFFFFF:
        raise ArchiveError("entity count exceeds serialized_size_t (uint32)")
    # each std::array is its count followed by its doubles
    arrays = geom.reshape(n * (2 if is_box else 1), dim)
    head = np.frombuffer(struct.pack(cnt, dim), dtype=np.uint8)
    rows = np.concatenate([np.broadcast_to(head, (arrays.shape[0], head.size)),
                           arrays.view(np.uint8).reshape(arrays.shape[0], 8 * dim)], axis=1)
    return struct.pack(cnt, n) + rows.tobytes()
