"""ctypes binding of libcqlhip.so (the C ABI declared in include/cqlhip.h).

Loading fails loudly when the library is not built: there is no CPU path
behind these entry points.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libcqlhip.so")

U8, U32, U64, S32, F64, B32 = C.c_uint8, C.c_uint32, C.c_uint64, C.c_int32, C.c_double, C.c_uint32
P = C.c_void_p


class QE_Trace(C.Structure):
    _fields_ = [
        ("path", U32),
        ("kernel_count", U32),
        ("rows_before", U64),
        ("rows_after", U64),
        ("rows_scanned", U64),
        ("bytes_read", U64),
        ("gpu_kernel_ns", U64),
        ("wall_ns", U64),
        ("compile_ns", U64),
        ("kernel", C.c_char * 64),
    ]

    def as_dict(self) -> dict:
        return {f: (getattr(self, f).decode() if f == "kernel" else getattr(self, f)) for f, _ in self._fields_}


MAX_GROUP_COLS = 4  # QE_AGG_MAX_GROUP_COLS


class QE_BatchQuery(C.Structure):
    _fields_ = [
        ("where", C.c_void_p),
        ("where_count", U32),
        ("group_count", U32),
        ("group_cols", C.c_int32 * MAX_GROUP_COLS),
        ("funcs", C.c_void_p),
        ("arg_cols", C.c_void_p),
        ("expr_count", U32),
        ("cap", U64),
        ("out_group_count", C.c_void_p),
        ("out_keys", C.c_void_p),
        ("out_cells", C.c_void_p),
        ("ok", U32),
        ("fused", U32),
    ]


class QE_BatchTrace(C.Structure):
    _fields_ = [(n, U32) for n in ("queries", "fused", "fallback", "launches", "items", "blocks_per_cu")] + [
        (n, U64) for n in ("rows_scanned", "bytes_read", "gpu_kernel_ns", "wall_ns")] + [
        (n, U32) for n in ("tuned_wg", "tuned_buffers")]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class QE_BatchConfig(C.Structure):
    _fields_ = [("wg", U32), ("items", U32), ("buffers", U32), ("blocks_per_cu", U32), ("lds_budget", U64)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class QE_JoinQuery(C.Structure):
    _fields_ = [
        ("dim", C.c_void_p),
        ("fact_key_col", U32),
        ("dim_key_col", U32),
        ("dim_where", C.c_void_p),
        ("dim_where_count", U32),
        ("group_count", U32),
        ("group_cols", C.c_int32 * MAX_GROUP_COLS),
        ("funcs", C.c_void_p),
        ("arg_cols", C.c_void_p),
        ("expr_count", U32),
        ("cap", U64),
        ("out_group_count", C.c_void_p),
        ("out_keys", C.c_void_p),
        ("out_cells", C.c_void_p),
        ("ok", U32),
    ]


class QE_JoinBatchTrace(C.Structure):
    _fields_ = [("queries", U32), ("runs", U32), ("fact_passes", U32), ("wall_ns", U64)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class QE_ScanQuery(C.Structure):
    _fields_ = [
        ("where", C.c_void_p),
        ("where_count", U32),
        ("count_only", U32),
        ("cap", U64),
        ("out_ids", C.c_void_p),
        ("out_count", C.c_void_p),
        ("ok", U32),
    ]


REQUEST_AGGREGATE, REQUEST_JOIN, REQUEST_SCAN = 0, 1, 2


class QE_ServeRequest(C.Structure):
    _fields_ = [("kind", U32), ("table", C.c_void_p), ("query", C.c_void_p)]


class QE_ServerConfig(C.Structure):
    _fields_ = [("workers", U32), ("max_batch", U32), ("hbm_budget", U64)]


class QE_ServerStats(C.Structure):
    _fields_ = [(n, U64) for n in ("submitted", "completed", "failed", "batches", "batched_queries", "p50_ns",
                                    "p99_ns", "max_ns", "mean_ns", "elapsed_ns", "inflight_peak_bytes",
                                    "queue_peak", "joins", "scans")]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class GPU_MemoryStats(C.Structure):
    _fields_ = [(n, U64) for n in ("bytes_in_use", "bytes_peak", "alloc_failures", "alloc_retry_successes",
                                    "pool_evictions", "pool_evicted_bytes", "alloc_calls", "unrecovered_alloc_failures")]


class QE_ResidentStats(C.Structure):
    _fields_ = [(n, U64) for n in ("hit_bytes", "miss_bytes", "evicted_bytes", "bytes_in_use", "entries",
                                    "budget_bytes")]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class QE_StreamStats(C.Structure):
    _fields_ = [(n, U64) for n in ("chunks", "chunks_skipped", "chunks_run", "chunks_resident", "transient_chunks",
                                    "bytes_copied", "copy_wait_ns", "wall_ns", "resident_whole")]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class String8(C.Structure):
    _fields_ = [("str", C.c_void_p), ("size", U64)]


class GPU_StreamStats(C.Structure):
    _fields_ = [("bytes_copied", U64), ("imports_created", U64), ("import_microseconds", U64), ("staged_bytes", U64),
                ("separate_copy_queue", B32)]


# (name, restype, argtypes) of every entry point declared in include/cqlhip.h
SIGNATURES = [
    ("gpu_hip_backend", P, []),
    ("gpu_hip_init", None, []),
    ("gpu_hip_release", None, []),
    ("gpu_hip_wait", None, []),
    ("gpu_hip_get_executed_kernel_time_microseconds", U64, []),
    ("gpu_hip_device_total_memory", U64, []),
    ("gpu_hip_device_free_memory", U64, []),
    ("gpu_hip_device_max_storage_buffer_range", U64, []),
    ("gpu_hip_device_lost", B32, []),
    ("gpu_hip_pool_begin_epoch", None, []),
    ("gpu_hip_memory_stats", GPU_MemoryStats, []),
    ("gpu_hip_buffer_alloc", P, [U64, U32, P]),
    ("gpu_hip_buffer_alloc_pooled", P, [String8, U64, U32, P]),
    ("gpu_hip_buffer_import_host_readonly", P, [P, U64]),
    ("gpu_hip_buffer_import_host_readonly_pooled", P, [String8, P, U64]),
    ("gpu_hip_buffer_release", None, [P]),
    ("gpu_hip_buffer_write", None, [P, P, U64]),
    ("gpu_hip_buffer_read", None, [P, P, U64]),
    ("gpu_hip_kernel_alloc", P, [String8]),
    ("gpu_hip_kernel_alloc_from_spirv", P, [String8, String8]),
    ("gpu_hip_kernel_alloc_from_spirv_spec", P, [String8, String8, P, U32]),
    ("gpu_hip_generated_kernel_stats", None, [P, P]),
    ("gpu_hip_kernel_release", None, [P]),
    ("gpu_hip_kernel_execute", None, [P, U32, U32]),
    ("gpu_hip_kernel_set_arg_buffer", None, [P, U32, P]),
    ("gpu_hip_kernel_set_arg_u64", None, [P, U32, U64]),
    ("gpu_hip_batch_begin", P, [U64, U64]),
    ("gpu_hip_batch_buffer_write", None, [P, P, P, U64]),
    ("gpu_hip_batch_buffer_zero", None, [P, P, U64]),
    ("gpu_hip_batch_buffer_fill", None, [P, P, U64, U32]),
    ("gpu_hip_batch_kernel_execute", None, [P, P, U32, U32]),
    ("gpu_hip_batch_buffer_read", None, [P, P, P, U64]),
    ("gpu_hip_batch_end", B32, [P]),
    ("gpu_hip_stream_alloc", P, [U32]),
    ("gpu_hip_stream_release", None, [P]),
    ("gpu_hip_stream_slot_begin", B32, [P, U32, U64, U64]),
    ("gpu_hip_stream_copy", B32, [P, U32, P, P, U64]),
    ("gpu_hip_stream_zero", None, [P, U32, P, U64]),
    ("gpu_hip_stream_fill", None, [P, U32, P, U64, U32]),
    ("gpu_hip_stream_dispatch", None, [P, U32, P, U32, U32]),
    ("gpu_hip_stream_read", None, [P, U32, P, P, U64]),
    ("gpu_hip_stream_submit", B32, [P, U32]),
    ("gpu_hip_stream_wait", B32, [P, U32]),
    ("gpu_hip_stream_get_stats", GPU_StreamStats, [P]),
    ("gpu_hip_scratch_arena", P, []),
    ("gpu_hip_name", String8, []),
    ("gpu_hip_compute_units", U32, []),
    ("gpu_hip_buffer_size", U64, [P]),
    ("gpu_hip_buffer_device_address", U64, [P]),
    ("gpu_hip_kernel_set_lds_bytes", None, [P, U32]),
    ("gpu_hip_max_lds_bytes", U32, []),
    ("gpu_hip_kernel_max_lds_bytes", U32, [P]),
    ("gpu_hip_kernel_clear_arg", None, [P, U32]),
    ("gpu_hip_kernel_set_arg_u32", None, [P, U32, U32]),
    ("gpu_hip_profile_enable", None, [B32]),
    ("gpu_hip_profile_reset", None, []),
    ("gpu_hip_profile_query", U64, [C.c_char_p, P]),
    ("gpu_hip_profile_tag", None, [C.c_char_p]),
    ("gpu_hip_batch_begin_graph", P, [C.c_char_p]),
    ("gpu_hip_graph_stats", None, [P, P]),
    ("gpu_hip_host_alloc", P, [U64]),
    ("gpu_hip_host_free", None, [P]),
    ("gpu_hip_set_evict_hook", None, [P]),
    ("gpu_hip_buffer_wrap_host", P, [P, U64]),
    ("gpu_hip_buffer_view", P, [P, U64, U64]),
    ("gpu_hip_fault_alloc_at", None, [U64]),
    ("gpu_hip_fault_read_at", None, [U64]),
    ("gpu_hip_buffer_write_checked", B32, [P, P, U64]),
    ("gpu_hip_buffer_read_checked", B32, [P, P, U64]),
    ("gpu_hip_last_error", C.c_char_p, []),
    ("qe_settings_set_u64", None, [C.c_char_p, U64]),
    ("qe_settings_u64", U64, [C.c_char_p, U64]),
    ("qe_settings_reset", None, []),
    ("qe_last_error", C.c_char_p, []),
    ("qe_table_alloc", P, [U64]),
    ("qe_table_release", None, [P]),
    ("qe_table_row_count", U64, [P]),
    ("qe_table_add_column", S32, [P, C.c_char_p, U32, P, P]),
    ("qe_table_column_buffer", P, [P, U32]),
    ("qe_table_column_read", B32, [P, U32, U64, U64, P]),
    ("qe_table_build_zonemap", B32, [P, U32]),
    ("qe_table_add_host_column", S32, [P, C.c_char_p, U32, P]),
    ("qe_resident_stats", None, [P]),
    ("qe_resident_clear", None, []),
    ("qe_stream_last_stats", None, [P]),
    ("qe_gen_events", B32, [P, U64, S32, S32, S32, S32]),
    ("qe_gen_scale_events", B32, [P, S32, S32, S32, S32]),
    ("qe_gen_fact_fk", B32, [P, U32, U64, P, U32]),
    ("qe_gen_dim", B32, [P, U64, U32, B32, S32, S32]),
    ("qe_scan_filter", B32, [P, P, U32, B32, P, U64, P, P]),
    ("qe_aggregate", B32, [P, P, U32, P, U32, P, P, U32, U64, P, P, P, P]),
    ("qe_aggregate_noalloc", B32, [P, P, U32, P, P, U32, P, P]),
    ("qe_aggregate_join", B32, [P, U32, P, U32, P, U32, P, U32, P, P, U32, U64, P, P, P, P]),
    ("qe_hash_join", U32, [P, U32, P, U64, P, U32, P, U64, U32, P, U64, P, P]),
    ("qe_gather", U32, [P, P, U32, P, U64, U64, U64, P]),
    ("qe_table_encode_column", S32, [P, U32, B32]),
    ("qe_scan_materialize", U32, [P, P, U32, P, U32, U64, P, P]),
    ("qe_aggregate_batch", B32, [P, P, U32, P]),
    ("qe_aggregate_batch_tune", B32, [P, P, U32, P, P]),
    ("qe_server_start", P, [P]),
    ("qe_server_submit", U64, [P, P, P]),
    ("qe_server_wait", B32, [P, U64, U64]),
    ("qe_server_stats", None, [P, P]),
    ("qe_server_stats_reset", None, [P]),
    ("qe_server_stop", None, [P]),
    ("qe_server_drive", B32, [P, P, P, U32, C.c_double, U64, U64, P]),
    ("qe_server_drive_mix", B32, [P, P, P, U32, C.c_double, U64, U64, P]),
    ("qe_server_submit_join", U64, [P, P, P]),
    ("qe_aggregate_join_batch", B32, [P, P, U32, P]),
    ("qe_server_submit_scan", U64, [P, P, P]),
    ("qe_server_drive_requests", B32, [P, P, U32, C.c_double, U64, U64, P]),
    ("qe_bandwidth_read", B32, [P, U64, U32, P, P]),
    ("qe_bandwidth_copy", B32, [P, P, U64, U32, P]),
]

_lib = None


class NativeMissing(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load libcqlhip.so; raises NativeMissing when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeMissing(f"{LIB_PATH} is not built: run `python -m computeql_amd.build` (no CPU fallback exists)")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    lib = load()
    e = lib.qe_last_error()
    return e.decode(errors="replace") if e else ""
