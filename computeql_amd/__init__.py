"""computeql_amd - MI355X (gfx950) HIP backend for ComputeQL's scan+filter ->
group-by/aggregate -> hash join path.

The product is the C-ABI library computeql_amd/lib/libcqlhip.so
(include/cqlhip.h); this package is its host-side Python mirror.
"""
from .engine import (AGG_PATHS, AVG, BUILD_NULL, COUNT, F32, F64, FUNC_NAMES, I32, I64, JOIN_TYPES, MAX, MIN, SCAN_PATHS,
                     SUM, U32, U64, AggResult, AggSpec, QueryError, Table, aggregate_batch, aggregate_join, autotune_batch,
                     bandwidth_copy, bandwidth_read, compute_units, get_setting, hash_join, profile, reset_settings, resident_clear,
                     resident_stats, set_setting, stream_stats)
from .ir import col, lit
from .native import LIB_PATH, NativeMissing, load

__all__ = [
    "AGG_PATHS", "AVG", "BUILD_NULL", "COUNT", "F32", "F64", "FUNC_NAMES", "I32", "I64", "JOIN_TYPES", "MAX", "MIN",
    "SCAN_PATHS", "SUM", "U32", "U64", "AggResult", "AggSpec", "QueryError", "Table", "aggregate_batch", "aggregate_join",
    "autotune_batch", "bandwidth_copy", "bandwidth_read", "compute_units", "get_setting", "hash_join",
    "profile", "reset_settings", "resident_clear", "resident_stats", "set_setting", "stream_stats", "col", "lit", "LIB_PATH", "NativeMissing", "load",
]
