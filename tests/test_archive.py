"""Static-core archive layout (serialization.h:359-386) on the CPU: the reader
against hand-derived archives, the hand-derived node tables against the oracle,
and the managed entity vector (serialization.h:433-439)."""
import struct

import numpy as np
import pytest

from octree_amd.archive import ArchiveError, entity_vector_bytes, read_static_archive
from oracle import oracle as O
from helpers import (HAND_BOXES_2D, HAND_POINTS_2D, archive_children, check_archive_vs_table,
                           hand_archive_box_2d, hand_archive_point_2d)


def test_hand_point_tree_matches_oracle():
    t = O.OracleTree(2, "point", np.array(HAND_POINTS_2D), [0, 0, 4, 4], 2, 1).export()
    a = read_static_archive(hand_archive_point_2d(), 2, False)
    assert a["storage"] == 0
    check_archive_vs_table(a, t, False, 1, [0, 0], [4, 4])
    assert archive_children(a["child_begins"], 4) == [(1, 4), (0, 0), (0, 0), (0, 0)]


def test_hand_box_tree_matches_oracle():
    t = O.OracleTree(2, "box", np.array(HAND_BOXES_2D), None, 2, 1).export()
    a = read_static_archive(hand_archive_box_2d(), 2, True)
    assert a["storage"] == 0
    check_archive_vs_table(a, t, True, 1, [0, 0], [4, 4])


def test_reader_rejects_damaged_archives():
    b = hand_archive_point_2d()
    with pytest.raises(ArchiveError):
        read_static_archive(b[:-1], 2, False)
    with pytest.raises(ArchiveError):
        read_static_archive(b + b"\0", 2, False)
    with pytest.raises(ArchiveError):  # read as a 3-D tree: arrays of 2 coordinates
        read_static_archive(b, 3, False)
    bad = bytearray(b)
    bad[0] = 1  # core SERIALIZED_VERSION_ID
    with pytest.raises(ArchiveError):
        read_static_archive(bytes(bad), 2, False)


def test_large_dataset_counts_are_u64():
    # the same point archive with serialized_size_t = uint64 (NodeStorage256
    # keeps its u8 fields): every count widens by 4 bytes
    small = hand_archive_point_2d()
    # m_nodeSize and its 4 arrays, variant index, begins, segments, depths,
    # entity ids, geometry and its 4 arrays
    counts = (1 + 4) + 1 + 1 + 1 + 1 + 1 + (1 + 4)
    b = struct.pack("<IIQI", 0, 0, 1, 2)
    b += struct.pack("<Q", 4) + b"".join(struct.pack("<Q2d", 2, s, s) for s in (4.0, 2.0, 1.0, 0.5))
    b += struct.pack("<Q", 0) + struct.pack("<Q", 2) + bytes([1, 4]) + struct.pack("<Q", 4) + bytes([0, 0, 0, 1, 1, 1, 2, 1])
    b += struct.pack("<Q", 4) + bytes([0, 2, 2, 2]) + struct.pack("<Q3I", 3, 0, 1, 2)
    b += struct.pack("<Q", 4) + b"".join(struct.pack("<Q2d", 2, x, y) for x, y in ((0, 0), (0, 0), (3, 0), (0, 3)))
    assert len(b) == len(small) + 4 * counts
    a, s = read_static_archive(b, 2, False, large_dataset=True), read_static_archive(small, 2, False)
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(s[k])), k


def test_managed_entity_vector():
    pts = np.array([[1.5, -2.0], [0.25, 8.0]])
    b = entity_vector_bytes(pts, 2, False)
    assert b == struct.pack("<I", 2) + struct.pack("<I2d", 2, 1.5, -2.0) + struct.pack("<I2d", 2, 0.25, 8.0)
    boxes = np.array([[0.0, 1.0, 2.0, 3.0]])
    b = entity_vector_bytes(boxes, 2, True, large_dataset=True)
    assert b == struct.pack("<Q", 1) + struct.pack("<Q2d", 2, 0.0, 1.0) + struct.pack("<Q2d", 2, 2.0, 3.0)
    full = hand_archive_box_2d() + entity_vector_bytes(np.array(HAND_BOXES_2D), 2, True)
    a = read_static_archive(full, 2, True, managed=True)
    assert np.array_equal(a["entities"], np.array(HAND_BOXES_2D))
