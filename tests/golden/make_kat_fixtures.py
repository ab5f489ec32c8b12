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
    "issue38_collision": {  # general.tests.cpp:4728-4772: 10 rounds x 1000 loose 3D boxes, mt19937(0); vs brute force
        "rounds": 10, "box_no": 1000, "space_max": 8.0, "size_max": 2.0, "depth": 10, "cap": 5,
        "src": "general.tests.cpp:4728",
    },
}

if __name__ == "__main__":
    out = Path(__file__).resolve().parent / "kat_reference.json"
    out.write_text(json.dumps(fixtures, indent=1))
    print(out)
