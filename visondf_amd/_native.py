"""ctypes binding of libvdf_hip.so (the C ABI declared in include/vdf.h).

The library is built in-tree (``python -c "import __graft_entry__ as g; g.build()"``
or ``make -C visondf_amd/csrc``).  There is no fallback: if the shared
library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvdf_hip.so")

# status codes
VDF_OK = 0
VDF_ERR_INVALID = 1
VDF_ERR_UNSUPPORTED = 2
VDF_ERR_OOM = 3
VDF_ERR_HIP = 4
VDF_ERR_BUDGET = 5
VDF_ERR_TIMEOUT = 6

# key types
KEY_U32, KEY_I32, KEY_U64, KEY_I64, KEY_CHAR16, KEY_F32, KEY_F64 = 0, 1, 2, 3, 4, 5, 6
KEY_BYTES = 7  # multi-column keys only: fixed-width byte strings
# value types
VAL_F64, VAL_I64, VAL_I32, VAL_U32, VAL_U64 = 0, 1, 2, 3, 4
# argsort flags
SORT_DESC, SORT_DESC_REVERSE, SORT_PERM64 = 0x1, 0x2, 0x4
# aggregates
AGG_COUNT, AGG_SUM, AGG_MEAN, AGG_MIN, AGG_MAX, AGG_ALL = 0x1, 0x2, 0x4, 0x8, 0x10, 0x1F

MAX_ROWS = (1 << 30) - 1
MAX_VALUE_COLS = 4
MAX_KEY_COLS = 4

# every symbol include/vdf.h declares
EXPORTS = (
    "vdf_last_error", "vdf_version", "vdf_create", "vdf_destroy",
    "vdf_argsort_scratch_bytes", "vdf_groupby_scratch_bytes", "vdf_groupby_batch_scratch_bytes",
    "vdf_argsort", "vdf_sort_by", "vdf_sort_by_noalloc", "vdf_gather", "vdf_groupby",
    "vdf_profile_enable", "vdf_profile_reset", "vdf_profile_read",
    "vdf_tune_set", "vdf_tune_get",
    "vdf_argsort_batch", "vdf_groupby_batch",
    "vdf_server_create", "vdf_server_destroy", "vdf_server_add_table", "vdf_server_submit",
    "vdf_server_wait", "vdf_server_release", "vdf_server_stats_get",
    "vdf_autotune_argsort", "vdf_autotune_sort_by", "vdf_autotune_groupby_batch",
    "vdf_plan_groupby_batch", "vdf_plan_launch", "vdf_plan_scratch_bytes", "vdf_plan_destroy",
    "vdf_groupby_multi", "vdf_argsort_multi", "vdf_argsort_bytes", "vdf_groupby_bytes", "vdf_broadcast",
    "vdf_grouping_create", "vdf_grouping_aggregate", "vdf_grouping_destroy",
    "vdf_read_csv", "vdf_csv_info", "vdf_csv_column_get", "vdf_csv_free",
)

CSV_STRING, CSV_CHAR, CSV_BOOL, CSV_INT, CSV_UINT, CSV_DOUBLE = 0, 1, 2, 3, 4, 5


class VdfCsvOptions(C.Structure):
    _fields_ = [("delim", C.c_char), ("quote", C.c_char), ("trailing", C.c_char), ("reserved0", C.c_char),
                ("header", C.c_int), ("int_bytes", C.c_int), ("char_width", C.c_int)]


class VdfCsvColumn(C.Structure):
    _fields_ = [("kind", C.c_int), ("width", C.c_uint32), ("data", C.c_void_p), ("name", C.c_char_p)]


class VdfValueCol(C.Structure):
    _fields_ = [("type", C.c_int), ("data", C.c_void_p)]


class VdfAggOut(C.Structure):
    _fields_ = [("sum", C.c_void_p), ("mean", C.c_void_p), ("min", C.c_void_p), ("max", C.c_void_p)]


class VdfGroupbyOut(C.Structure):
    _fields_ = [
        ("n_groups", C.c_void_p),
        ("group_keys", C.c_void_p),
        ("offsets", C.c_void_p),
        ("counts", C.c_void_p),
        ("perm", C.c_void_p),
        ("aggs", VdfAggOut * MAX_VALUE_COLS),
        ("capacity", C.c_uint64),
    ]


class VdfServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("workers", C.c_int), ("max_batch_tables", C.c_int), ("reserved", C.c_int),
                ("arena_bytes", C.c_uint64), ("max_batch_rows", C.c_uint64)]


class VdfServed(C.Structure):
    _fields_ = [("status", C.c_int), ("batch_tables", C.c_uint32), ("n_groups", C.c_uint64),
                ("submit_ns", C.c_uint64), ("start_ns", C.c_uint64), ("done_ns", C.c_uint64),
                ("out", VdfGroupbyOut)]


class VdfServerStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("submitted", "completed", "failed", "launches", "retries", "rows",
                                          "queued", "arena_bytes", "arena_used", "arena_peak")]


class VdfKeyCol(C.Structure):
    _fields_ = [("type", C.c_int), ("data", C.c_void_p), ("width", C.c_uint32)]


class VdfTuneReport(C.Structure):
    _fields_ = [("trials", C.c_int), ("sort_items", C.c_int), ("sort_blocks_per_cu", C.c_int), ("sort_xcd", C.c_int),
                ("sort_chunk", C.c_int), ("hist_rows", C.c_int), ("digit_bits", C.c_int), ("sort_local", C.c_int),
                ("default_ms", C.c_double), ("best_ms", C.c_double)]


class VdfSortSeg(C.Structure):
    _fields_ = [("keys", C.c_void_p), ("n", C.c_uint64), ("perm_out", C.c_void_p), ("sorted_keys_out", C.c_void_p)]


class VdfTable(C.Structure):
    _fields_ = [("keys", C.c_void_p), ("n", C.c_uint64), ("vals", C.c_void_p * MAX_VALUE_COLS)]


class VdfError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"vdf status {status}: {msg}")
        self.status = status


_lib = None


def load() -> C.CDLL:
    """Load libvdf_hip.so once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    vp, u64, i32, u32 = C.c_void_p, C.c_uint64, C.c_int, C.c_uint32
    lib.vdf_last_error.restype = C.c_char_p
    lib.vdf_last_error.argtypes = []
    lib.vdf_version.restype = C.c_char_p
    lib.vdf_version.argtypes = []
    lib.vdf_create.restype = i32
    lib.vdf_create.argtypes = [C.POINTER(vp), i32, u64]
    lib.vdf_destroy.restype = i32
    lib.vdf_destroy.argtypes = [vp]
    lib.vdf_argsort_scratch_bytes.restype = u64
    lib.vdf_argsort_scratch_bytes.argtypes = [i32, u64, u32, i32]
    lib.vdf_read_csv.restype = i32
    lib.vdf_read_csv.argtypes = [vp, vp, u64, C.POINTER(VdfCsvOptions), C.POINTER(vp)]
    lib.vdf_csv_info.restype = i32
    lib.vdf_csv_info.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_int)]
    lib.vdf_csv_column_get.restype = i32
    lib.vdf_csv_column_get.argtypes = [vp, i32, C.POINTER(VdfCsvColumn)]
    lib.vdf_csv_free.restype = i32
    lib.vdf_csv_free.argtypes = [vp]
    lib.vdf_groupby_scratch_bytes.restype = u64
    lib.vdf_groupby_scratch_bytes.argtypes = [i32, u64]
    lib.vdf_groupby_batch_scratch_bytes.restype = u64
    lib.vdf_groupby_batch_scratch_bytes.argtypes = [i32, C.POINTER(C.c_uint64), i32, C.POINTER(C.c_int), i32]
    lib.vdf_argsort.restype = i32
    lib.vdf_argsort.argtypes = [vp, i32, vp, u64, u32, vp, vp, vp]
    lib.vdf_sort_by.restype = i32
    lib.vdf_sort_by.argtypes = [vp, i32, vp, u64, u32, vp, vp, u32, vp, vp, vp]
    lib.vdf_sort_by_noalloc.restype = i32
    lib.vdf_sort_by_noalloc.argtypes = [vp, i32, vp, u64, u32, vp, vp, u32, vp, vp, vp]
    lib.vdf_gather.restype = i32
    lib.vdf_gather.argtypes = [vp, u32, vp, vp, i32, u64, vp, vp]
    lib.vdf_groupby.restype = i32
    lib.vdf_groupby.argtypes = [vp, i32, vp, u64, C.POINTER(VdfValueCol), i32, u32,
                                C.POINTER(VdfGroupbyOut), vp]
    lib.vdf_profile_enable.restype = i32
    lib.vdf_profile_enable.argtypes = [vp, i32]
    lib.vdf_profile_reset.restype = i32
    lib.vdf_profile_reset.argtypes = [vp]
    lib.vdf_profile_read.restype = i32
    lib.vdf_profile_read.argtypes = [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(u64), C.POINTER(u64)]
    lib.vdf_argsort_batch.restype = i32
    lib.vdf_argsort_batch.argtypes = [vp, i32, C.POINTER(VdfSortSeg), i32, u32, vp]
    lib.vdf_groupby_batch.restype = i32
    lib.vdf_groupby_batch.argtypes = [vp, i32, C.POINTER(VdfTable), i32, C.POINTER(C.c_int), i32, u32,
                                      C.POINTER(VdfGroupbyOut), vp]
    lib.vdf_server_create.restype = i32
    lib.vdf_server_create.argtypes = [C.POINTER(vp), C.POINTER(VdfServerConfig)]
    lib.vdf_server_destroy.restype = i32
    lib.vdf_server_destroy.argtypes = [vp]
    lib.vdf_server_add_table.restype = i32
    lib.vdf_server_add_table.argtypes = [vp, i32, vp, u64, C.POINTER(C.c_int), C.POINTER(vp), i32, u64,
                                         C.POINTER(u32)]
    lib.vdf_server_submit.restype = i32
    lib.vdf_server_submit.argtypes = [vp, u32, u32, i32, C.POINTER(u64)]
    lib.vdf_server_wait.restype = i32
    lib.vdf_server_wait.argtypes = [vp, u64, C.c_int64, C.POINTER(VdfServed)]
    lib.vdf_server_release.restype = i32
    lib.vdf_server_release.argtypes = [vp, u64]
    lib.vdf_server_stats_get.restype = i32
    lib.vdf_server_stats_get.argtypes = [vp, C.POINTER(VdfServerStats)]
    lib.vdf_autotune_argsort.restype = i32
    lib.vdf_autotune_argsort.argtypes = [vp, i32, vp, u64, u32, vp, i32, vp, C.POINTER(VdfTuneReport)]
    lib.vdf_autotune_sort_by.restype = i32
    lib.vdf_autotune_sort_by.argtypes = [vp, i32, vp, u64, u32, vp, u32, vp, vp, i32, vp, C.POINTER(VdfTuneReport)]
    lib.vdf_autotune_groupby_batch.restype = i32
    lib.vdf_autotune_groupby_batch.argtypes = [vp, i32, C.POINTER(VdfTable), i32, C.POINTER(C.c_int), i32, u32,
                                               C.POINTER(VdfGroupbyOut), i32, vp, C.POINTER(VdfTuneReport)]
    lib.vdf_plan_groupby_batch.restype = i32
    lib.vdf_plan_groupby_batch.argtypes = [vp, i32, C.POINTER(VdfTable), i32, C.POINTER(C.c_int), i32, u32,
                                           C.POINTER(vp)]
    lib.vdf_plan_launch.restype = i32
    lib.vdf_plan_launch.argtypes = [vp, C.POINTER(VdfTable), i32, C.POINTER(VdfGroupbyOut), vp]
    lib.vdf_plan_scratch_bytes.restype = i32
    lib.vdf_plan_scratch_bytes.argtypes = [vp, C.POINTER(u64)]
    lib.vdf_plan_destroy.restype = i32
    lib.vdf_plan_destroy.argtypes = [vp]
    lib.vdf_groupby_multi.restype = i32
    lib.vdf_groupby_multi.argtypes = [vp, C.POINTER(VdfKeyCol), i32, u64, C.POINTER(VdfValueCol), i32, u32,
                                      C.POINTER(VdfGroupbyOut), C.POINTER(vp), vp]
    lib.vdf_argsort_multi.restype = i32
    lib.vdf_argsort_multi.argtypes = [vp, C.POINTER(VdfKeyCol), i32, u64, u32, vp, vp]
    lib.vdf_argsort_bytes.restype = i32
    lib.vdf_argsort_bytes.argtypes = [vp, vp, u32, u64, u32, vp, vp, vp]
    lib.vdf_groupby_bytes.restype = i32
    lib.vdf_groupby_bytes.argtypes = [vp, vp, u32, u64, C.POINTER(VdfValueCol), i32, u32, C.POINTER(VdfGroupbyOut), vp]
    lib.vdf_broadcast.restype = i32
    lib.vdf_broadcast.argtypes = [vp, i32, i32, vp, vp, vp, u64, vp, u64, vp, vp]
    lib.vdf_grouping_create.restype = i32
    lib.vdf_grouping_create.argtypes = [vp, i32, vp, u64, vp, C.POINTER(vp)]
    lib.vdf_grouping_aggregate.restype = i32
    lib.vdf_grouping_aggregate.argtypes = [vp, C.POINTER(VdfValueCol), i32, u32, C.POINTER(VdfGroupbyOut), vp]
    lib.vdf_grouping_destroy.restype = i32
    lib.vdf_grouping_destroy.argtypes = [vp]
    lib.vdf_tune_set.restype = i32
    lib.vdf_tune_set.argtypes = [vp, C.c_char_p, C.c_int64]
    lib.vdf_tune_get.restype = i32
    lib.vdf_tune_get.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != VDF_OK:
        raise VdfError(status, load().vdf_last_error().decode())
