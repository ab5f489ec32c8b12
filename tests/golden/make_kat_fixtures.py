"""Writes tests/golden/kat_reference.json: known-answer vectors transcribed
from the reference's own unit tests (tests/unittests/general.tests.cpp), with
the file:line each one comes from.  Inputs and expected outputs only; the
doubles that the tests build from expressions (sqrt2 * 0.5, -1 / sqrt(5)) are
evaluated here with IEEE double arithmetic, as the C++ tests do."""
import json
import math
from pathlib import Path

SQ2H = 1.4142135623730951 * 0.5  # std::numbers::sqrt2 * 0.5

point_set_no1 = [  # getPointSetNo1<2>() general.tests.cpp:72-130
    [0.0, 0.0], [3.0, 3.0], [5.0, 3.0], [5.2, 3.0], [7.0, 1.0], [0.0, 5.0], [1.0, 5.0], [0.0, 7.0], [1.0, 7.0],
    [3.0, 7.0], [3.0, 7.0], [2.2, 4.5], [2.5, 4.5], [3.4, 4.5], [3.6, 4.6], [2.4, 5.5], [2.5, 5.5], [3.5, 5.5],
    [3.5, 5.5], [5.0, 5.0], [5.0, 5.0], [7.0, 5.0], [7.0, 5.0], [6.5, 6.5], [6.5, 6.5], [7.5, 6.5], [7.5, 6.5],
    [6.5, 7.5], [6.5, 7.5], [7.5, 7.5], [7.5, 7.5], [8.0, 8.0],
]

boxes4 = [[0.0, 0.0, 1.0, 1.0], [1.0, 1.0, 2.0, 2.0], [2.0, 2.0, 3.0, 3.0], [3.0, 3.0, 4.0, 4.0]]
boxes5 = boxes4 + [[1.2, 1.2, 2.8, 2.8]]
boxes_r5 = [[2.0, 0.0, 3.0, 1.0], [3.0, 1.0, 4.0, 2.0], [3.0, 2.0, 4.0, 3.0], [2.0, 3.0, 4.0, 4.0], [2.5, 2.5, 3.5, 3.5]]
EPS = 2.220446049250313e-16  # std::numeric_limits<double>::epsilon()
INV_SQRT3 = float("0.577350269189625764509148780502")  # std::numbers::inv_sqrt3 (correctly rounded)
INV_SQRT5 = 1.0 / math.sqrt(5.0)  # 1.0 / std::sqrt(5.0)


SQRT2 = math.sqrt(2.0)  # std::sqrt(2.0)
F32_001 = 0.009999999776482582  # 0.01f widened to double

# PlaneSearch_3D_* general.tests.cpp:2407-2470
_PLANE_PTS_A = [[2.0, -1.0, 0.0], [3.0, -2.0, 1.0], [4.0, -3.0, 0.02], [4.0, -2.0, 0.0], [5.0, 0.0, 0.00999]]
_PLANE_PTS_B = _PLANE_PTS_A[:4] + [[5.0, 0.0, 1.00999]]
# PointGeneral3D adapter.tests.cpp:194-198
_PTS10 = [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [2.0, 2.0, 2.0], [3.0, 3.0, 3.0], [4.0, 4.0, 4.0], [0.0, 0.0, 4.0],
          [0.0, 4.0, 0.0], [4.0, 0.0, 0.0], [1.5, 1.5, 1.0], [1.0, 1.5, 1.5]]


def _plane_boxes(z3lo, z3hi, z4lo, z4hi):
    """The five boxes of the PlaneIntersection_3D_* tests (general.tests.cpp:2966-3075); boxes 3 and 4 vary in z."""
    return [[0.0, 0.0, 0.0, 1.0, 1.0, 0.0], [1.0, 1.0, -1.0, 2.0, 2.0, 0.0], [2.0, 2.0, -1.0, 3.0, 3.0, -0.2],
            [3.0, 3.0, z3lo, 4.0, 4.0, z3hi], [1.2, 1.2, z4lo, 2.8, 2.8, z4hi]]


def lift_1d(boxes):
    """1-D boxes [lo, hi] of a reference test as 2-D boxes with the second axis [0, 1] for every box.

    Every box spans the added axis, so a point with 0.5 there lies inside each box on that axis and
    any two boxes overlap strictly on it: the 1-D test's expected id set is the 2-D answer."""
    return [[lo, 0.0, hi, 1.0] for lo, hi in boxes]


# PickSearch_Issue8 general.tests.cpp:3776-4365: 144 cells {x0, 0, z0} - {x0 + 63, 0, z0 + 63}
_ISSUE8_CELLS = [  # (x0, z0) of mGridCells[0..143] in order
    (-1024, 1024), (-1024, 1088), (-1024, 1152), (-1024, 1216), (-1024, 1280), (-1024, 1344), (-1024, 1408), (-1024, 1472),
    (-960, 1024), (-960, 1088), (-960, 1152), (-960, 1216), (-960, 1280), (-960, 1344), (-960, 1408), (-960, 1472),
    (-896, 1024), (-896, 1088), (-896, 1152), (-896, 1216), (-896, 1280), (-896, 1344), (-896, 1408), (-896, 1472),
    (-832, 1024), (-832, 1088), (-832, 1152), (-832, 1216), (-832, 1280), (-832, 1344), (-832, 1408), (-832, 1472),
    (-768, 1024), (-768, 1088), (-768, 1152), (-768, 1216), (-768, 1280), (-768, 1344), (-768, 1408), (-768, 1472),
    (-704, 1024), (-704, 1088), (-704, 1152), (-704, 1216), (-704, 1280), (-704, 1344), (-704, 1408), (-704, 1472),
    (-640, 1024), (-640, 1088), (-640, 1152), (-640, 1216), (-640, 1280), (-640, 1344), (-640, 1408), (-640, 1472),
    (-576, 1024), (-576, 1088), (-576, 1152), (-576, 1216), (-576, 1280), (-576, 1344), (-576, 1408), (-576, 1472),
    (-1152, 896), (-1152, 960), (-1152, 1024), (-1152, 1088), (-1152, 1152), (-1152, 1216), (-1152, 1280), (-1152, 1344),
    (-1152, 1408), (-1152, 1472), (-1152, 1536), (-1152, 1600), (-1088, 896), (-1088, 960), (-1088, 1024), (-1088, 1088),
    (-1088, 1152), (-1088, 1216), (-1088, 1280), (-1088, 1344), (-1088, 1408), (-1088, 1472), (-1088, 1536), (-1088, 1600),
    (-512, 896), (-512, 960), (-512, 1024), (-512, 1088), (-512, 1152), (-512, 1216), (-512, 1280), (-512, 1344),
    (-512, 1408), (-512, 1472), (-512, 1536), (-512, 1600), (-448, 896), (-448, 960), (-448, 1024), (-448, 1088),
    (-448, 1152), (-448, 1216), (-448, 1280), (-448, 1344), (-448, 1408), (-448, 1472), (-448, 1536), (-448, 1600),
    (-1024, 896), (-1024, 960), (-960, 896), (-960, 960), (-896, 896), (-896, 960), (-832, 896), (-832, 960),
    (-768, 896), (-768, 960), (-704, 896), (-704, 960), (-640, 896), (-640, 960), (-576, 896), (-576, 960),
    (-1024, 1536), (-1024, 1600), (-960, 1536), (-960, 1600), (-896, 1536), (-896, 1600), (-832, 1536), (-832, 1600),
    (-768, 1536), (-768, 1600), (-704, 1536), (-704, 1600), (-640, 1536), (-640, 1600), (-576, 1536), (-576, 1600),
]

fixtures = {
    "morton_encode": [  # general.tests.cpp:136-217 (MortonTest)
        {"dim": 2, "grid": [0, 0], "code": 0, "src": "general.tests.cpp:152"},
        {"dim": 2, "grid": [2, 0], "code": 4, "src": "general.tests.cpp:158"},
        {"dim": 2, "grid": [0, 2], "code": 8, "src": "general.tests.cpp:164"},
        {"dim": 2, "grid": [2, 2], "code": 12, "src": "general.tests.cpp:170"},
        {"dim": 2, "grid": [1, 3], "code": 11, "src": "general.tests.cpp:176"},
        {"dim": 3, "grid": [0, 0, 0], "code": 0, "src": "general.tests.cpp:182"},
        {"dim": 3, "grid": [1, 0, 0], "code": 1, "src": "general.tests.cpp:188"},
        {"dim": 3, "grid": [0, 0, 1], "code": 4, "src": "general.tests.cpp:194"},
        {"dim": 3, "grid": [1, 1, 1], "code": 7, "src": "general.tests.cpp:200"},
        {"dim": 4, "grid": [1, 1, 1, 1], "code": 15, "src": "general.tests.cpp:206"},
        {"dim": 4, "grid": [2, 1, 1, 1], "code": 30, "src": "general.tests.cpp:212"},
        {"dim": 1, "grid": [0], "code": 0, "src": "general.tests.cpp:139"},
        {"dim": 1, "grid": [4], "code": 4, "src": "general.tests.cpp:145"},
    ],
    "node_key_1d": [  # general.tests.cpp:465-494 (GetNodeID / GetDepthID / RemoveSentinelBit, D=1)
        {"loc": 0, "depth": 0, "key": 1, "src": "general.tests.cpp:465"},
        {"loc": 1, "depth": 1, "key": 3, "src": "general.tests.cpp:470"},
        {"loc": 2, "depth": 2, "key": 6, "src": "general.tests.cpp:474"},
        {"loc": 7, "depth": 3, "key": 15, "src": "general.tests.cpp:479"},
    ],
    "knn_point_set_no1": {  # Point_kNNTest general.tests.cpp:2475-2602, tree (points, 3, nullopt, 3)
        "points": point_set_no1, "dim": 2, "max_depth": 3, "cap": 3,
        "cases": [
            {"q": [3.5, 5.5], "k": 2, "tol": 0.0, "expect": [17, 18], "src": "general.tests.cpp:2477"},
            {"q": [3.5, 5.5], "k": 3, "expect": [17, 18, 14], "src": "general.tests.cpp:2488"},
            {"q": [3.5, 5.5], "k": 4, "expect": [17, 18, 14, 16], "src": "general.tests.cpp:2499"},
            {"q": [3.5, 5.5], "k": 4, "max_within": 0.0, "expect": [17, 18], "src": "general.tests.cpp:2510"},
            {"q": [3.5, 5.5], "k": 5, "max_within": 1.0, "expect": [17, 18, 14, 16], "src": "general.tests.cpp:2521"},
            {"q": [3.5, 5.5], "k": 100, "expect": list(range(32)), "src": "general.tests.cpp:2532"},
            {"q": [2.0, 2.0], "k": 1, "expect": [1], "src": "general.tests.cpp:2547"},
            {"q": [2.0, 2.0], "k": 2, "expect": [1, 11], "src": "general.tests.cpp:2558"},
            {"q": [2.0, 2.0], "k": 3, "expect": [1, 11, 12], "src": "general.tests.cpp:2569"},
            {"q": [-1.0, -1.0], "k": 1, "expect": [0], "src": "general.tests.cpp:2582"},
            {"q": [-1.0, -1.0], "k": 5, "expect": [0, 1, 5, 6, 11], "src": "general.tests.cpp:2593"},
        ],
    },
    "knn_issue9": [  # Issue9_2D / Issue9_6D general.tests.cpp:2661-2765: nearest = brute-force argmin
        {"dim": 2, "max_depth": 9, "box": [0.0, 0.0, 100.0, 100.0], "cap": 16, "src": "general.tests.cpp:2674",
         "points": [[78.2619, 77.843], [90.3005, 90.5172], [69.8652, 12.2467], [48.4675, 48.4948], [36.3226, 68.4619],
                    [98.8799, 42.7149], [31.412, 38.6866], [63.2748, 77.0524], [62.8844, 17.0536], [80.8931, 39.8099],
                    [77.426, 64.9844], [81.9552, 25.009], [87.6088, 51.319], [78.5609, 80.4623], [51.3967, 39.5269],
                    [32.2042, 81.8779], [79.1739, 81.5467], [95.2619, 40.4742], [86.437, 92.4406], [3.95388, 60.2327],
                    [31.1283, 44.4917], [35.6778, 79.8545], [50.9926, 66.1373], [3.16271, 65.2519], [56.3665, 45.3819]],
         "q": [43.6406, 57.5691]},
        {"dim": 6, "max_depth": 9, "box": [0.0] * 6 + [100.0] * 6, "cap": 16, "src": "general.tests.cpp:2724",
         "points": [[50.2232, 0.276687, 37.7662, 41.2776, 26.3818, 74.0284], [35.8946, 83.7503, 97.1127, 47.2895, 40.9232, 83.7666],
                    [33.3541, 63.0669, 3.86075, 47.7923, 19.8039, 87.5608], [89.0684, 67.3278, 50.867, 49.8193, 72.6692, 54.0271],
                    [75.7099, 53.1241, 39.624, 40.4669, 13.6433, 88.6247], [43.7641, 93.7985, 68.8286, 9.71882, 2.16644, 12.1925],
                    [82.5195, 30.6964, 71.0556, 48.4744, 99.2295, 70.4137], [28.4568, 37.8366, 74.8597, 27.897, 60.6816, 0.247821],
                    [58.3617, 17.0165, 15.1021, 70.9832, 22.5325, 34.8085], [33.1004, 72.6729, 35.7043, 35.2888, 94.9917, 17.652],
                    [80.9693, 7.41406, 2.18394, 40.2606, 1.63274, 65.8996], [88.4851, 73.8366, 55.0264, 77.6467, 61.6751, 33.7444],
                    [48.1777, 90.1285, 75.9069, 49.1867, 39.7186, 74.2862], [68.0826, 73.0622, 7.85933, 60.8879, 41.3229, 16.6928],
                    [19.1693, 49.1172, 91.055, 52.5548, 16.2326, 0.610075], [19.3085, 81.9363, 55.8924, 1.77161, 29.4124, 72.3873],
                    [57.8801, 28.3918, 45.7062, 92.1774, 0.742745, 29.0499], [74.2952, 84.6691, 46.3357, 76.7894, 68.8394, 24.1707],
                    [0.199418, 93.2845, 94.9628, 38.2295, 82.1998, 59.2499], [68.4689, 0.261607, 75.5001, 58.427, 55.4243, 18.7392],
                    [53.9164, 95.4966, 59.657, 71.0292, 82.4362, 53.9452]],
         "q": [78.8658, 64.0361, 18.7755, 61.4618, 14.3312, 40.0196]},
    ],
    "ray_first_2d": [  # general.tests.cpp:1280-1520, QuadtreeBox(boxes, depth, nullopt, cap)
        {"boxes": boxes4, "depth": 3, "cap": 2, "o": [0.0, -1.0], "d": [1.0, 0.0], "tol": 0.0, "expect": [], "src": "general.tests.cpp:1280"},
        {"boxes": boxes4, "depth": 3, "cap": 2, "o": [5.0, 0.0], "d": [0.0, 1.0], "tol": 0.0, "expect": [], "src": "general.tests.cpp:1296"},
        {"boxes": boxes4, "depth": 3, "cap": 2, "o": [4.2, 4.4], "d": [SQ2H, SQ2H], "tol": 0.0, "expect": [], "src": "general.tests.cpp:1312"},
        {"boxes": boxes4, "depth": 3, "cap": 2, "o": [1.5, 2.6], "d": [-SQ2H, -SQ2H], "tol": 0.0, "expect": [], "src": "general.tests.cpp:1328"},
        {"boxes": boxes4, "depth": 3, "cap": 2, "o": [7.0, 2.0], "d": [-SQ2H, -SQ2H], "tol": 0.0, "expect": [], "src": "general.tests.cpp:1344"},
        {"boxes": boxes4, "depth": 3, "cap": 2, "o": [0.0, 0.5], "d": [1.0, 0.0], "tol": 0.0, "expect": [0], "src": "general.tests.cpp:1360"},
        {"boxes": boxes4, "depth": 3, "cap": 2, "o": [0.0, 2.0], "d": [1.0, 0.0], "tol": 0.0, "expect": [2, 1], "src": "general.tests.cpp:1377"},
        {"boxes": boxes4, "depth": 3, "cap": 2, "o": [1.5, 1.0], "d": [1.0, 0.0], "tol": 0.0, "expect": [1], "src": "general.tests.cpp:1395"},
        {"boxes": boxes4, "depth": 3, "cap": 2, "o": [1.5, 1.0], "d": [-1.0, 0.0], "tol": 0.0, "expect": [0, 1], "src": "general.tests.cpp:1412"},
        {"boxes": boxes4, "depth": 3, "cap": 2, "o": [3.5, 3.5], "d": [-1.0 / math.sqrt(5.0), -2.0 / math.sqrt(5.0)], "tol": 0.0,
         "expect": [3], "src": "general.tests.cpp:1430"},
        {"boxes": [[0.0, 0.0, 5.0, 1.0], [1.0, 0.0, 3.0, 1.0], [2.0, 0.0, 4.0, 1.0]], "depth": 3, "cap": 2, "o": [-1.0, 0.5],
         "d": [1.0, 0.0], "tol": 0.0, "expect": [2, 1, 0], "src": "general.tests.cpp:1446"},
        {"boxes": [[1.0, 0.0, 2.0, 1.0]], "depth": 1, "cap": 1, "o": [0.0, 1.001], "d": [1.0, 0.0], "tol": 0.01,
         "expect": [0], "src": "general.tests.cpp:1465"},
        {"boxes": [[0.0, 0.0, 1.0, 1.0], [10.0, 0.0, 11.0, 1.0]], "depth": 3, "cap": 2, "o": [-1.0, 0.5], "d": [1.0, 0.0],
         "tol": 0.0, "inc": 0.5, "expect": [0], "src": "general.tests.cpp:1482"},
        {"boxes": [[0.0, 0.0, 3.0, 1.0], [1.0, 0.0, 2.0, 1.0], [10.0, 0.0, 11.0, 1.0]], "depth": 3, "cap": 2, "o": [-1.0, 0.5],
         "d": [1.0, 0.0], "tol": 0.0, "inc": 0.0, "max_dist": 5.0, "expect": [1, 0], "src": "general.tests.cpp:1500"},
    ],
    "knn_box_2d": [  # box-tree kNN (MinMaxDistance) general.tests.cpp; expected sets (order not asserted there)
        {"boxes": [[0.20, 0.00, 2.31, 2.31], [1.00, 0.00, 1.01, 1.41], [0.40, 0.00, 0.89, 0.89], [2.10, 0.00, 3.66, 3.66],
                   [-2.21, 0.00, -0.20, 2.21], [-2.50, 0.00, -0.90, 2.47]],
         "depth": 3, "cap": 2, "q": [0.0, 0.0], "k": 1, "tol": 2.220446049250313e-16, "expect": [0, 2, 4],
         "src": "general.tests.cpp:2648"},
        {"boxes": [[0.0, 0.0, 1.0, 1.0], [1.0, 1.0, 3.0, 3.0], [4.0, 4.0, 5.0, 5.0]], "depth": 3, "cap": 2,
         "q": [2.0, 2.0], "k": 1, "expect": [1], "src": "general.tests.cpp:2772"},
        {"boxes": [[0.0, 0.0, 1.0, 1.0], [1.0, 1.0, 3.0, 3.0], [4.0, 4.0, 5.0, 5.0]], "depth": 3, "cap": 2,
         "q": [-1.0, 1.0], "k": 1, "expect": [0], "src": "general.tests.cpp:2788"},
    ],
    "pick": [  # PickSearch KATs; expected sets (the tests compare permutations)
        {"dim": 2, "boxes": boxes5, "depth": 3, "cap": 2, "p": [2.5, 2.5], "expect": [2, 4], "src": "general.tests.cpp:2928"},
        {"dim": 2, "boxes": boxes5, "depth": 3, "cap": 2, "p": [2.0, 2.0], "expect": [1, 2, 4], "src": "general.tests.cpp:2947"},
        {"dim": 2, "boxes": lift_1d([[0, 1], [3, 4], [4, 5], [5, 6], [7, 8]]), "depth": 3, "cap": 2, "p": [4.0, 0.5],
         "expect": [1, 2], "src": "general.tests.cpp:3672 (1-D, lifted)"},
        {"dim": 2, "boxes": lift_1d([[0, 16], [0, 1], [3, 4], [4, 5], [5, 6], [2, 4], [4, 6], [2, 6]]), "depth": 4, "cap": 2,
         "p": [4.0, 0.5], "expect": [0, 2, 3, 5, 6, 7], "src": "general.tests.cpp:3694 (1-D, lifted)"},
        {"dim": 2, "boxes": lift_1d([[-8, 8], [0, 1], [-4, -3], [4, 5], [-6, -5], [-4, 2], [4, 6], [-6, 2]]), "depth": 4,
         "cap": 2, "p": [-4.0, 0.5], "expect": [0, 2, 5, 7], "src": "general.tests.cpp:3721 (1-D, lifted)"},
        {"dim": 2, "boxes": lift_1d([[-17, -1], [-2, -1], [-4, -3], [-5, -4], [-6, -5], [-4, -2], [-6, -4], [-6, -2]]),
         "depth": 4, "cap": 2, "p": [-4.0, 0.5], "expect": [0, 2, 3, 5, 6, 7], "src": "general.tests.cpp:3748 (1-D, lifted)"},
        {"dim": 3, "boxes": [[x, 0.0, z, x + 63.0, 0.0, z + 63.0] for x, z in _ISSUE8_CELLS], "depth": 3, "cap": 21,
         "p": [-864.0, 0.0, 1471.0], "tol": 0.0, "expect": [22], "src": "general.tests.cpp:3776"},
        {"dim": 3, "points": [[1.0, 1.0, 1.0], [2.0, 2.0, 2.0], [3.0, 3.0, 3.0]], "depth": 3, "cap": 2,
         "p": [2.0, 2.0, 2.0], "tol": 0.1, "expect": [1], "src": "general.tests.cpp:5293"},
    ],
    "ray_all": [  # RayIntersectedAll KATs (tolerance given as the tests give it); "ordered": the test asserts the order
        {"dim": 2, "boxes": boxes_r5, "depth": 3, "cap": 2, "o": [2.7, 2.6], "d": [SQ2H, SQ2H], "tol": 0.0,
         "expect": [4, 2, 3], "ordered": True, "src": "general.tests.cpp:1545"},
        {"dim": 2, "boxes": boxes_r5, "depth": 3, "cap": 2, "o": [2.6, 2.8], "d": [SQ2H, SQ2H], "tol": 0.0,
         "expect": [3, 4], "src": "general.tests.cpp:1563"},
        {"dim": 3, "boxes": [[2.0, 0.0, 1.0, 3.0, 1.0, 2.0], [3.0, 1.0, 1.0, 4.0, 2.0, 2.0], [3.0, 2.0, 1.0, 4.0, 3.0, 2.0],
                             [2.0, 3.0, 1.5, 4.0, 4.0, 2.0], [2.5, 2.5, 1.5, 3.5, 3.5, 2.0]],
         "depth": 3, "cap": 2, "o": [2.6, 2.8, 1.0], "d": [INV_SQRT3] * 3, "tol": 0.0, "expect": [3, 4],
         "src": "general.tests.cpp:1581"},
        {"dim": 5, "boxes": [[2.0, 0.0, 1.0, 0.0, 0.0, 3.0, 1.0, 2.0, 2.0, 2.0], [3.0, 1.0, 1.0, 0.0, 0.0, 4.0, 2.0, 2.0, 2.0, 2.0],
                             [3.0, 2.0, 1.0, 0.0, 0.0, 4.0, 3.0, 2.0, 2.0, 2.0], [2.0, 3.0, 1.5, 0.0, 0.0, 4.0, 4.0, 2.0, 2.0, 2.0],
                             [2.5, 2.5, 1.5, 0.0, 0.0, 3.5, 3.5, 2.0, 2.0, 2.0]],
         "depth": 3, "cap": 2, "o": [2.6, 2.8, 1.0, 0.0, 0.0], "d": [INV_SQRT5] * 5, "tol": 0.0, "expect": [3, 4],
         "src": "general.tests.cpp:1599"},
        {"dim": 2, "boxes": boxes5, "depth": 3, "cap": 2, "o": [1.1, 2.0], "d": [1.0, 0.0], "tol": 0.0,
         "expect": [1, 4, 2], "src": "general.tests.cpp:1639"},
        {"dim": 2, "boxes": boxes5, "depth": 3, "cap": 2, "o": [0.0, 2.001], "d": [1.0, 0.0], "tol": 0.001 - EPS,
         "expect": [4, 2], "src": "general.tests.cpp:1645"},
        {"dim": 2, "boxes": boxes5, "depth": 3, "cap": 2, "o": [2.0005, 2.0], "d": [1.0, 0.0], "tol": 0.001,
         "expect": [2, 4], "src": "general.tests.cpp:1651"},
        {"dim": 2, "boxes": boxes5, "depth": 3, "cap": 2, "o": [1.0, 1.9], "d": [1.0, 0.0], "tol": 0.899,
         "expect": [1, 4, 2], "src": "general.tests.cpp:1657"},
        {"dim": 2, "boxes": boxes5, "depth": 3, "cap": 2, "o": [1.0, 1.9], "d": [1.0, 0.0], "tol": 0.9 - EPS,
         "expect": [1, 4, 2], "src": "general.tests.cpp:1663"},
        {"dim": 2, "boxes": boxes5, "depth": 3, "cap": 2, "o": [1.0, 1.9], "d": [1.0, 0.0], "tol": 0.901,
         "expect": [0, 1, 4, 2], "src": "general.tests.cpp:1669"},
        {"dim": 2, "boxes": [[1.0, 0.01, 2.0, 0.5]], "depth": 3, "cap": 2, "o": [0.0, 0.0], "d": [1.0, 0.0], "tol": 0.0,
         "inc": 0.0, "expect": [], "src": "general.tests.cpp:1697"},
        {"dim": 2, "boxes": [[1.0, 0.01, 2.0, 0.5]], "depth": 3, "cap": 2, "o": [0.0, 0.0], "d": [1.0, 0.0], "tol": 0.0,
         "inc": 0.01, "expect": [0], "src": "general.tests.cpp:1708"},
        {"dim": 2, "boxes": [[2.0, -1.0, 3.0, 1.0], [10.0, -1.0, 11.0, 1.0]], "depth": 3, "cap": 2, "o": [0.0, 0.0],
         "d": [1.0, 0.0], "tol": 0.0, "inc": 0.0, "max_dist": 5.0, "expect": [0], "src": "general.tests.cpp:1733"},
        {"dim": 2, "boxes": [[2.0, -1.0, 3.0, 1.0], [10.0, -1.0, 11.0, 1.0]], "depth": 3, "cap": 2, "o": [0.0, 0.0],
         "d": [1.0, 0.0], "tol": 0.0, "inc": 0.0, "max_dist": 100.0, "expect": [0, 1], "ordered": True,
         "src": "general.tests.cpp:1746"},
    ],
    "collision_two_trees": [  # static CollisionDetection(leftTree, left, rightTree, right); expected (left, right) sets
        {"left": boxes4, "right": boxes_r5, "depth": 3, "cap": 2, "expect": [[3, 3], [2, 4], [3, 4]],
         "src": "general.tests.cpp:3081"},
        {"left": lift_1d([[0.0, 1.0], [1.0, 2.0], [2.0, 3.0], [3.0, 4.0]]), "right": lift_1d([[3.5, 4.5], [4.5, 5.5], [5.5, 6.5]]),
         "depth": 3, "cap": 1, "expect": [[3, 0]], "src": "general.tests.cpp:2202 (1-D, lifted)"},
        {"left": lift_1d([[0.0, 4.0], [0.0, 2.0], [2.0, 4.0], [0.0, 1.0], [1.0, 2.0], [2.0, 3.0], [3.0, 4.0]]),
         "right": lift_1d([[2.0, 6.0], [2.0, 4.0], [4.0, 6.0], [2.0, 3.0], [3.0, 4.0], [4.0, 5.0], [5.0, 6.0]]),
         "depth": 3, "cap": 1,
         "expect": [[0, 0], [0, 1], [2, 0], [2, 1], [0, 3], [2, 3], [5, 0], [5, 1], [5, 3], [0, 4], [2, 4], [6, 0], [6, 1],
                    [6, 4]],
         "src": "general.tests.cpp:2235 (1-D, lifted)"},
    ],
    "planes": [  # PlaneSearch / PlaneIntersection / FrustumCulling KATs; expected sets (tests compare permutations)
        # planes: [origo distance, normal...] each; mode 0 PlaneSearch, 2 FrustumCulling
        {"kind": "point", "dim": 3, "geom": _PLANE_PTS_A, "depth": 3, "cap": 2, "mode": 0,
         "planes": [[0.0, 0.0, 0.0, 1.0]], "tol": 0.01, "expect": [0, 3, 4], "src": "general.tests.cpp:2419"},
        {"kind": "point", "dim": 3, "geom": _PLANE_PTS_B, "depth": 3, "cap": 2, "mode": 0,
         "planes": [[1.0, 0.0, 0.0, 1.0]], "tol": 0.01, "expect": [1, 4], "src": "general.tests.cpp:2436"},
        {"kind": "point", "dim": 3, "geom": _PLANE_PTS_B, "depth": 3, "cap": 2, "mode": 0,
         "planes": [[4.0, 1.0, 0.0, 0.0]], "tol": 0.01, "expect": [2, 3], "src": "general.tests.cpp:2453"},
        {"kind": "box", "dim": 3, "geom": _plane_boxes(1.0, 1.2, -1.0, 1.0), "depth": 3, "cap": 2, "mode": 0,
         "planes": [[0.0, 0.0, 0.0, 1.0]], "tol": 0.01, "expect": [0, 1, 4], "src": "general.tests.cpp:2977"},
        {"kind": "box", "dim": 3, "geom": _plane_boxes(0.8, 0.9, -1.0, 1.1), "depth": 3, "cap": 2, "mode": 0,
         "planes": [[1.0, 0.0, 0.0, 1.0]], "tol": 0.01, "expect": [4], "src": "general.tests.cpp:2994"},
        {"kind": "box", "dim": 3, "geom": _plane_boxes(1.0, 1.2, -1.1, 1.1), "depth": 3, "cap": 2, "mode": 0,
         "planes": [[-1.0, 0.0, 0.0, 1.0]], "tol": 0.01, "expect": [1, 2, 4], "src": "general.tests.cpp:3010"},
        {"kind": "box", "dim": 3, "geom": _plane_boxes(1.0, 1.2, -1.1, 1.1), "depth": 3, "cap": 2, "mode": 0,
         "planes": [[1.0, 1.0, 0.0, 0.0]], "tol": 0.01, "expect": [0, 1], "src": "general.tests.cpp:3027"},
        {"kind": "box", "dim": 3, "geom": _plane_boxes(1.0, 1.2, -1.1, 1.1), "depth": 3, "cap": 2, "mode": 0,
         "planes": [[-1.0, 1.0, 0.0, 0.0]], "tol": 0.01, "expect": [], "src": "general.tests.cpp:3043"},
        {"kind": "box", "dim": 3, "geom": _plane_boxes(1.0, 1.2, -1.1, 1.1), "depth": 3, "cap": 2, "mode": 0,
         "planes": [[SQRT2, 0.0, SQRT2 * 0.5, SQRT2 * 0.5]], "tol": 0.01, "expect": [1, 2, 4],
         "src": "general.tests.cpp:3075"},
        {"kind": "point", "dim": 3, "geom": _PTS10, "depth": 3, "cap": 1, "mode": 0,
         "planes": [[2.6, 1.0 / math.sqrt(3.0), 1.0 / math.sqrt(3.0), 1.0 / math.sqrt(3.0)]], "tol": 0.3,
         "expect": [5, 6, 7, 8, 9], "src": "adapter.tests.cpp:208"},
        {"kind": "point", "dim": 3, "geom": _PTS10, "depth": 3, "cap": 1, "mode": 2,
         "planes": [[2 * SQRT2, 1.0 / SQRT2, 1.0 / SQRT2, 0.0], [2 * SQRT2, 0.0, 1.0 / SQRT2, 1.0 / SQRT2],
                    [0.0, 1.0 / SQRT2, -1.0 / SQRT2, 0.0]], "tol": 0.01, "expect": [2, 3, 4],
         "src": "adapter.tests.cpp:217 (Eigen planes: origo distance = -offset)"},
        {"kind": "box", "dim": 2, "geom": boxes5, "depth": 3, "cap": 1, "mode": 2,
         "planes": [[2 * SQRT2, 1.0 / SQRT2, 1.0 / SQRT2], [-2.0, 0.0, -1.0]], "tol": F32_001, "expect": [1, 2, 4],
         "src": "adapter.tests.cpp:283 (tolerance 0.01f)"},
    ],
    "issue38_collision": {  # general.tests.cpp:4728-4772: 10 rounds x 1000 loose 3D boxes, mt19937(0); vs brute force
        "rounds": 10, "box_no": 1000, "space_max": 8.0, "size_max": 2.0, "depth": 10, "cap": 5,
        "src": "general.tests.cpp:4728",
    },
}

if __name__ == "__main__":
    out = Path(__file__).resolve().parent / "kat_reference.json"
    out.write_text(json.dumps(fixtures, indent=1))
    print(out)
