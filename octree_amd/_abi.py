"""ctypes binding of libotg's C ABI (include/otg/otg.h).

The library must be built in-tree (python -m octree_amd.build); importing this
module without it raises immediately -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libotg.so"

OTG_OK = 0
OTG_ERR_INVALID_ARG = 1
OTG_ERR_HIP = 2
OTG_ERR_OOM = 3
OTG_ERR_UNSUPPORTED = 4
OTG_ERR_OUT_OF_SPACE = 5
OTG_ERR_KEY_OVERFLOW = 6
OTG_ERR_NO_DEVICE = 7
OTG_ERR_CAPACITY = 8
OTG_MEM_HOST = 0
OTG_MEM_DEVICE = 1
OTG_KIND_POINT = 0
OTG_KIND_BOX = 1
OTG_RANGE_INSIDE = 0
OTG_RANGE_OVERLAP = 1
OTG_PLANE_SEARCH = 0
OTG_PLANE_POSITIVE = 1
OTG_PLANE_FRUSTUM = 2
OTG_ARCHIVE_LARGE_DATASET = 1

STATUS_NAMES = {
    0: "OTG_OK",
    1: "OTG_ERR_INVALID_ARG",
    2: "OTG_ERR_HIP",
    3: "OTG_ERR_OOM",
    4: "OTG_ERR_UNSUPPORTED",
    5: "OTG_ERR_OUT_OF_SPACE",
    6: "OTG_ERR_KEY_OVERFLOW",
    7: "OTG_ERR_NO_DEVICE",
    8: "OTG_ERR_CAPACITY",
}

# every symbol declared in include/otg/otg.h
EXPORTED = [
    "otg_ctx_create", "otg_ctx_destroy", "otg_ctx_synchronize", "otg_status_string", "otg_ctx_last_error",
    "otg_ctx_mem_stats", "otg_ctx_release_cache", "otg_ctx_set_profiling", "otg_ctx_kernel_times",
    "otg_ctx_reset_kernel_times", "otg_tree_build", "otg_tree_build_seq", "otg_forest_build",
    "otg_tree_deserialize", "otg_tree_free",
    "otg_tree_get_info",
    "otg_tree_get_set_info", "otg_tree_export", "otg_tree_serialize", "otg_knn", "otg_range", "otg_ray_first", "otg_collide", "otg_knn_footprint",
    "otg_collide_trees", "otg_pick", "otg_ray_all", "otg_planes", "otg_knn_noalloc",
    "otg_knn_plan_create", "otg_knn_plan_run", "otg_knn_plan_result", "otg_knn_plan_copy", "otg_knn_plan_free",
    "otg_build_plan_create", "otg_build_plan_run", "otg_build_plan_tree", "otg_build_plan_free",
    "otg_step_plan_create", "otg_step_plan_run", "otg_step_plan_free",
    "otg_range_plan_create", "otg_range_plan_run", "otg_range_plan_result", "otg_range_plan_copy",
    "otg_range_plan_free",
    "otg_autotune", "otg_ctx_set_tuning", "otg_ctx_get_tuning",
    "otg_result_get_info", "otg_result_copy", "otg_result_device", "otg_result_free",
    "otg_server_create", "otg_server_destroy", "otg_server_add_tree", "otg_server_remove_tree",
    "otg_server_has_tree", "otg_server_submit", "otg_server_wait", "otg_server_wait_detail", "otg_server_stats",
    "otg_server_set_batch_limit",
]


class TreeInfo(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("dim", C.c_int32), ("nsets", C.c_uint32), ("reserved", C.c_uint32),
        ("entity_count", C.c_uint64), ("node_count", C.c_uint64), ("device_bytes", C.c_uint64),
    ]


class SetInfo(C.Structure):
    _fields_ = [
        ("entity_offset", C.c_uint64), ("entity_count", C.c_uint64), ("node_offset", C.c_uint64),
        ("node_count", C.c_uint64), ("max_depth", C.c_uint32), ("max_elements", C.c_uint32),
        ("space_min", C.c_double * 16), ("space_max", C.c_double * 16),
    ]


class Tuning(C.Structure):
    _fields_ = [("leaf_capacity", C.c_uint32), ("index_capacity", C.c_uint32), ("waves_per_simd", C.c_uint32),
                ("order_queries", C.c_uint32), ("persistent", C.c_uint32), ("range_records", C.c_uint32),
                ("step_ms", C.c_double)]


class OtgError(RuntimeError):
    def __init__(self, status: int, what: str, detail: str = ""):
        self.status = status
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)}" + (f" ({detail})" if detail else ""))


_lib = None

P = C.c_void_p
PP = C.POINTER(C.c_void_p)
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)
F64P = C.POINTER(C.c_double)
U8P = C.POINTER(C.c_uint8)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m octree_amd.build` (no CPU fallback exists)")
    L = C.CDLL(str(LIB_PATH))
    sig = {
        "otg_ctx_create": (C.c_int, [C.c_int, PP]),
        "otg_ctx_destroy": (None, [P]),
        "otg_ctx_synchronize": (C.c_int, [P]),
        "otg_status_string": (C.c_char_p, [C.c_int]),
        "otg_ctx_last_error": (C.c_char_p, [P]),
        "otg_ctx_mem_stats": (C.c_int, [P, U64P, U64P]),
        "otg_ctx_release_cache": (C.c_int, [P]),
        "otg_ctx_set_profiling": (C.c_int, [P, C.c_int]),
        "otg_ctx_kernel_times": (C.c_int, [P, C.c_uint32, C.c_char_p, F64P, U64P, U32P]),
        "otg_ctx_reset_kernel_times": (C.c_int, [P]),
        "otg_tree_build": (C.c_int, [P, C.c_int, C.c_int, C.c_uint64, P, C.c_int, P, C.c_uint32, C.c_uint32, PP]),
        "otg_forest_build": (C.c_int, [P, C.c_int, C.c_int, C.c_uint32, U64P, P, C.c_int, P, P, U32P, PP]),
        "otg_tree_free": (None, [P]),
        "otg_tree_get_info": (C.c_int, [P, C.POINTER(TreeInfo)]),
        "otg_tree_get_set_info": (C.c_int, [P, C.c_uint32, C.POINTER(SetInfo)]),
        "otg_tree_export": (C.c_int, [P, C.c_uint32, P, P, P, P, P, P, P, P, P, P]),
        "otg_tree_serialize": (C.c_int, [P, C.c_uint32, C.c_uint32, P, C.c_uint64, U64P]),
        "otg_knn": (C.c_int, [P, C.c_uint64, P, P, C.c_int, C.c_uint32, C.c_double, C.c_double, PP]),
        "otg_autotune": (C.c_int, [P, C.c_int, C.c_int, C.c_uint32, P, P, C.c_int, P, P, C.c_uint64, P, P, C.c_uint32,
                                   C.c_uint64, P, P, C.c_int, C.POINTER(Tuning)]),
        "otg_ctx_set_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
        "otg_ctx_get_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
        "otg_knn_noalloc": (C.c_int, [P, C.c_uint64, P, P, C.c_uint32, C.c_double, C.c_double, P, P, C.c_uint64,
                                      C.POINTER(C.c_uint64)]),
        "otg_tree_build_seq": (C.c_int, [P, C.c_int, C.c_int, C.c_uint64, P, C.c_int, P, C.c_uint32, C.c_uint32, PP]),
        "otg_tree_deserialize": (C.c_int, [P, C.c_int, C.c_int, P, C.c_uint64, C.c_uint32, C.c_uint64, P, C.c_int,
                                           PP]),
        "otg_knn_plan_create": (C.c_int, [P, C.c_uint64, P, P, C.c_uint32, C.c_double, C.c_double, PP]),
        "otg_knn_plan_run": (C.c_int, [P]),
        "otg_knn_plan_result": (C.c_int, [P, PP, PP, C.POINTER(C.c_uint64)]),
        "otg_knn_plan_copy": (C.c_int, [P, P, P]),
        "otg_knn_plan_free": (None, [P]),
        "otg_build_plan_create": (C.c_int, [P, C.c_int, C.c_uint32, U64P, P, P, U32P, U32P, PP]),
        "otg_build_plan_run": (C.c_int, [P]),
        "otg_build_plan_tree": (C.c_int, [P, PP]),
        "otg_build_plan_free": (None, [P]),
        "otg_step_plan_create": (C.c_int, [P, P, P, PP]),
        "otg_range_plan_create": (C.c_int, [P, C.c_uint64, P, P, C.c_int, C.c_double, PP]),
        "otg_range_plan_run": (C.c_int, [P]),
        "otg_range_plan_result": (C.c_int, [P, PP, PP, C.POINTER(C.c_uint64)]),
        "otg_range_plan_copy": (C.c_int, [P, P, P]),
        "otg_range_plan_free": (None, [P]),
        "otg_step_plan_run": (C.c_int, [P]),
        "otg_step_plan_free": (None, [P]),
        "otg_knn_footprint": (C.c_int, [P, C.c_uint64, P, P, C.c_int, C.c_uint32, C.c_double, C.c_double, U64P]),
        "otg_range": (C.c_int, [P, C.c_uint64, P, P, C.c_int, C.c_int, C.c_double, PP]),
        "otg_ray_first": (C.c_int, [P, C.c_uint64, P, P, C.c_int, C.c_double, C.c_double, C.c_double, PP]),
        "otg_collide": (C.c_int, [P, C.c_uint32, C.c_double, PP]),
        "otg_collide_trees": (C.c_int, [P, C.c_uint32, P, C.c_uint32, C.c_double, PP]),
        "otg_pick": (C.c_int, [P, C.c_uint64, P, P, C.c_int, C.c_double, PP]),
        "otg_planes": (C.c_int, [P, C.c_uint64, C.c_uint32, P, P, C.c_int, C.c_int, C.c_double, PP]),
        "otg_ray_all": (C.c_int, [P, C.c_uint64, P, P, C.c_int, C.c_double, C.c_double, C.c_double, PP]),
        "otg_result_get_info": (C.c_int, [P, U64P, U64P]),
        "otg_result_copy": (C.c_int, [P, P, P]),
        "otg_result_device": (C.c_int, [P, PP, PP]),
        "otg_result_free": (None, [P]),
        "otg_server_create": (C.c_int, [C.c_int, C.c_uint64, C.c_uint32, PP]),
        "otg_server_destroy": (None, [P]),
        "otg_server_add_tree": (C.c_int, [P, C.c_uint32, C.c_int, C.c_int, C.c_uint64, P, C.c_int, P, C.c_uint32,
                                          C.c_uint32]),
        "otg_server_remove_tree": (C.c_int, [P, C.c_uint32]),
        "otg_server_has_tree": (C.c_int, [P, C.c_uint32, C.POINTER(C.c_int)]),
        "otg_server_submit": (C.c_int, [P, C.c_uint32, C.c_int, C.c_uint64, P, C.c_uint32, C.c_double, C.c_double,
                                        C.c_double, C.c_int, U64P]),
        "otg_server_wait": (C.c_int, [P, C.c_uint64, PP, F64P, F64P]),
        "otg_server_wait_detail": (C.c_int, [P, C.c_uint64, PP, C.POINTER(C.c_double)]),
        "otg_server_stats": (C.c_int, [P, U64P]),
        "otg_server_set_batch_limit": (C.c_int, [P, C.c_uint64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
