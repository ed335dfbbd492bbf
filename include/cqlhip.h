/*
 * cqlhip.h - C ABI of the MI355X (gfx950) HIP execution backend for
 * ComputeQL's scan+filter -> group-by/aggregate -> hash join path.
 *
 * Two layers, both plain C (no HIP or torch types in any signature):
 *
 *  1. The GPU_Backend function table (gpu_backend.h:17-71 of the reference)
 *     implemented over HIP: buffers, library kernels, run-time compiled
 *     kernels, batches and stream slots.  A ComputeQL build registers it as
 *     GPU_BackendKind_HIP (gpu_backend.h:8, gpu_inc.h:6-19) with
 *       gpu_backend_register(GPU_BackendKind_HIP, str8_lit("hip"), gpu_hip_backend());
 *     and selects it with --gpu=hip (main.c:135-139).
 *
 *  2. The operator entry points of query_exec.c for this path (scan/filter,
 *     global and grouped aggregates, join+aggregate), driving (1) the same
 *     way the reference's host code drives its Vulkan backend.  Each cites
 *     the reference function it replaces.
 *
 * All reference citations are file:line in the ComputeQL source tree.
 */
#ifndef CQLHIP_H
#define CQLHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* base types of base_core.h; a ComputeQL translation unit that already has
 * them (and gpu.h / gpu_backend.h) defines CQLHIP_IN_COMPUTEQL first */
#ifndef CQLHIP_IN_COMPUTEQL
typedef uint8_t  U8;
typedef uint16_t U16;
typedef uint32_t U32;
typedef uint64_t U64;
typedef int32_t  S32;
typedef int64_t  S64;
typedef uint32_t B32;
typedef float    F32;
typedef double   F64;

typedef struct String8 { U8* str; U64 size; } String8;

/* ------------------------------------------------------------------ */
/* 1. GPU_Backend (gpu.h:10-159, gpu_backend.h:17-71)                  */
/* ------------------------------------------------------------------ */
typedef enum GPU_BufferFlags {
  GPU_BufferFlag_Read = (1 << 0),
  GPU_BufferFlag_Write = (1 << 1),
  GPU_BufferFlag_ReadWrite = (1 << 2),
  GPU_BufferFlag_HostVisible = (1 << 3),
  GPU_BufferFlag_DeviceLocal = (1 << 4),
  GPU_BufferFlag_HostCached = (1 << 5),
  GPU_BufferFlag_ZeroCopy = (1 << 6),
  GPU_BufferFlag_CopyHostPointer = (1 << 7),
  GPU_BufferFlag_Dynamic = (1 << 8),
  GPU_BufferFlag_Shared = (1 << 9),
} GPU_BufferFlags;

typedef struct GPU_MemoryStats { /* gpu.h:25-36 */
  U64 bytes_in_use;
  U64 bytes_peak;
  U64 alloc_failures;
  U64 alloc_retry_successes;
  U64 pool_evictions;
  U64 pool_evicted_bytes;
  U64 alloc_calls;
  U64 unrecovered_alloc_failures;
} GPU_MemoryStats;

typedef struct GPU_StreamStats { /* gpu.h:104-111 */
  U64 bytes_copied;
  U64 imports_created;
  U64 import_microseconds;
  U64 staged_bytes;
  B32 separate_copy_queue;
} GPU_StreamStats;

typedef struct GPU_Buffer GPU_Buffer;
typedef struct GPU_Kernel GPU_Kernel;
typedef struct GPU_Batch GPU_Batch;
typedef struct GPU_Stream GPU_Stream;

/* Same member order as struct GPU_Backend, gpu_backend.h:17-71.
 * scratch_arena returns the backend's host scratch arena (opaque here). */
typedef struct GPU_Backend {
  U32 struct_size;
  void (*init)(void);
  void (*release)(void);
  void (*wait)(void);
  U64 (*get_executed_kernel_time_microseconds)(void);
  U64 (*device_total_memory)(void);
  U64 (*device_free_memory)(void);
  U64 (*device_max_storage_buffer_range)(void);
  B32 (*device_lost)(void);
  void (*pool_begin_epoch)(void);
  GPU_MemoryStats (*memory_stats)(void);
  GPU_Buffer* (*buffer_alloc)(U64 size, GPU_BufferFlags flags, void* data);
  GPU_Buffer* (*buffer_alloc_pooled)(String8 name, U64 size, GPU_BufferFlags flags, void* data);
  GPU_Buffer* (*buffer_import_host_readonly)(void* host_ptr, U64 size);
  GPU_Buffer* (*buffer_import_host_readonly_pooled)(String8 name, void* host_ptr, U64 size);
  void (*buffer_release)(GPU_Buffer* buffer);
  void (*buffer_write)(GPU_Buffer* buffer, void* data, U64 size);
  void (*buffer_read)(GPU_Buffer* buffer, void* data, U64 size);
  GPU_Kernel* (*kernel_alloc)(String8 name);
  GPU_Kernel* (*kernel_alloc_from_spirv)(String8 name, String8 spirv);
  GPU_Kernel* (*kernel_alloc_from_spirv_spec)(String8 name, String8 spirv, U32* spec_values, U32 spec_count);
  void (*generated_kernel_stats)(U64* out_compiles, U64* out_evictions);
  void (*kernel_release)(GPU_Kernel* kernel);
  void (*kernel_execute)(GPU_Kernel* kernel, U32 global_work_size, U32 local_work_size);
  void (*kernel_set_arg_buffer)(GPU_Kernel* kernel, U32 index, GPU_Buffer* buffer);
  void (*kernel_set_arg_u64)(GPU_Kernel* kernel, U32 index, U64 value);
  GPU_Batch* (*batch_begin)(U64 upload_bytes_needed, U64 download_bytes_needed);
  void (*batch_buffer_write)(GPU_Batch* batch, GPU_Buffer* buffer, void* data, U64 size);
  void (*batch_buffer_zero)(GPU_Batch* batch, GPU_Buffer* buffer, U64 size);
  void (*batch_buffer_fill)(GPU_Batch* batch, GPU_Buffer* buffer, U64 size, U32 value);
  void (*batch_kernel_execute)(GPU_Batch* batch, GPU_Kernel* kernel, U32 global_work_size, U32 local_work_size);
  void (*batch_buffer_read)(GPU_Batch* batch, GPU_Buffer* buffer, void* out_data, U64 size);
  B32 (*batch_end)(GPU_Batch* batch);
  GPU_Stream* (*stream_alloc)(U32 slot_count);
  void (*stream_release)(GPU_Stream* stream);
  B32 (*stream_slot_begin)(GPU_Stream* stream, U32 slot, U64 upload_bytes, U64 readback_bytes);
  B32 (*stream_copy)(GPU_Stream* stream, U32 slot, GPU_Buffer* destination, void* host_ptr, U64 size);
  void (*stream_zero)(GPU_Stream* stream, U32 slot, GPU_Buffer* buffer, U64 size);
  void (*stream_fill)(GPU_Stream* stream, U32 slot, GPU_Buffer* buffer, U64 size, U32 value);
  void (*stream_dispatch)(GPU_Stream* stream, U32 slot, GPU_Kernel* kernel, U32 global_work_size, U32 local_work_size);
  void (*stream_read)(GPU_Stream* stream, U32 slot, GPU_Buffer* buffer, void* out_data, U64 size);
  B32 (*stream_submit)(GPU_Stream* stream, U32 slot);
  B32 (*stream_wait)(GPU_Stream* stream, U32 slot);
  GPU_StreamStats (*stream_get_stats)(GPU_Stream* stream);
  void* (*scratch_arena)(void);
  String8 (*name)(void);
} GPU_Backend;
#endif /* CQLHIP_IN_COMPUTEQL */

/* The filled table; replaces the Vulkan registration of gpu_dispatch.c:1-28. */
GPU_Backend* gpu_hip_backend(void);

/* The individual entries (gpu_vulkan.c counterparts named in the .cpp). */
void gpu_hip_init(void);
void gpu_hip_release(void);
void gpu_hip_wait(void);
U64 gpu_hip_get_executed_kernel_time_microseconds(void);
U64 gpu_hip_device_total_memory(void);
U64 gpu_hip_device_free_memory(void);
U64 gpu_hip_device_max_storage_buffer_range(void);
B32 gpu_hip_device_lost(void);
void gpu_hip_pool_begin_epoch(void);
GPU_MemoryStats gpu_hip_memory_stats(void);
GPU_Buffer* gpu_hip_buffer_alloc(U64 size, GPU_BufferFlags flags, void* data);
GPU_Buffer* gpu_hip_buffer_alloc_pooled(String8 name, U64 size, GPU_BufferFlags flags, void* data);
GPU_Buffer* gpu_hip_buffer_import_host_readonly(void* host_ptr, U64 size);
GPU_Buffer* gpu_hip_buffer_import_host_readonly_pooled(String8 name, void* host_ptr, U64 size);
void gpu_hip_buffer_release(GPU_Buffer* buffer);
void gpu_hip_buffer_write(GPU_Buffer* buffer, void* data, U64 size);
void gpu_hip_buffer_read(GPU_Buffer* buffer, void* data, U64 size);
GPU_Kernel* gpu_hip_kernel_alloc(String8 name);
GPU_Kernel* gpu_hip_kernel_alloc_from_spirv(String8 name, String8 code_object);
GPU_Kernel* gpu_hip_kernel_alloc_from_spirv_spec(String8 name, String8 code_object, U32* spec_values, U32 spec_count);
void gpu_hip_generated_kernel_stats(U64* out_compiles, U64* out_evictions);
void gpu_hip_kernel_release(GPU_Kernel* kernel);
void gpu_hip_kernel_execute(GPU_Kernel* kernel, U32 global_work_size, U32 local_work_size);
void gpu_hip_kernel_set_arg_buffer(GPU_Kernel* kernel, U32 index, GPU_Buffer* buffer);
void gpu_hip_kernel_set_arg_u64(GPU_Kernel* kernel, U32 index, U64 value);
GPU_Batch* gpu_hip_batch_begin(U64 upload_bytes_needed, U64 download_bytes_needed);
void gpu_hip_batch_buffer_write(GPU_Batch* batch, GPU_Buffer* buffer, void* data, U64 size);
void gpu_hip_batch_buffer_zero(GPU_Batch* batch, GPU_Buffer* buffer, U64 size);
void gpu_hip_batch_buffer_fill(GPU_Batch* batch, GPU_Buffer* buffer, U64 size, U32 value);
void gpu_hip_batch_kernel_execute(GPU_Batch* batch, GPU_Kernel* kernel, U32 global_work_size, U32 local_work_size);
void gpu_hip_batch_buffer_read(GPU_Batch* batch, GPU_Buffer* buffer, void* out_data, U64 size);
B32 gpu_hip_batch_end(GPU_Batch* batch);
GPU_Stream* gpu_hip_stream_alloc(U32 slot_count);
void gpu_hip_stream_release(GPU_Stream* stream);
B32 gpu_hip_stream_slot_begin(GPU_Stream* stream, U32 slot, U64 upload_bytes, U64 readback_bytes);
B32 gpu_hip_stream_copy(GPU_Stream* stream, U32 slot, GPU_Buffer* destination, void* host_ptr, U64 size);
void gpu_hip_stream_zero(GPU_Stream* stream, U32 slot, GPU_Buffer* buffer, U64 size);
void gpu_hip_stream_fill(GPU_Stream* stream, U32 slot, GPU_Buffer* buffer, U64 size, U32 value);
void gpu_hip_stream_dispatch(GPU_Stream* stream, U32 slot, GPU_Kernel* kernel, U32 global_work_size, U32 local_work_size);
void gpu_hip_stream_read(GPU_Stream* stream, U32 slot, GPU_Buffer* buffer, void* out_data, U64 size);
B32 gpu_hip_stream_submit(GPU_Stream* stream, U32 slot);
B32 gpu_hip_stream_wait(GPU_Stream* stream, U32 slot);
GPU_StreamStats gpu_hip_stream_get_stats(GPU_Stream* stream);
void* gpu_hip_scratch_arena(void);
String8 gpu_hip_name(void);

/* HIP-specific extensions (no Vulkan counterpart). */
U32 gpu_hip_compute_units(void);          /* CUs of the device (256 on MI355X) */
U64 gpu_hip_buffer_size(GPU_Buffer* b);   /* bytes of a buffer */
U64 gpu_hip_buffer_device_address(GPU_Buffer* b); /* device pointer as an integer */
/* dynamic LDS bytes for the next dispatches of a kernel (the Vulkan side
 * sizes shared memory with specialization constants instead) */
void gpu_hip_kernel_set_lds_bytes(GPU_Kernel* kernel, U32 bytes);
/* leave a binding intentionally empty (address 0) without the missing-arg
 * mark that binding a null buffer sets */
void gpu_hip_kernel_clear_arg(GPU_Kernel* kernel, U32 index);
/* set one 32-bit push word (the half of U64 slot word/2) */
void gpu_hip_kernel_set_arg_u32(GPU_Kernel* kernel, U32 word, U32 value);
U32 gpu_hip_max_lds_bytes(void);          /* LDS a workgroup may use */
/* the dynamic LDS one kernel's workgroup may take: gpu_hip_max_lds_bytes
 * less the kernel's own static __shared__ bytes; a dispatch asking for
 * more fails with an error instead of reaching the queue */
U32 gpu_hip_kernel_max_lds_bytes(GPU_Kernel* kernel);
/* kernel-time profiling: hipEvents around every dispatch of `name` (a
 * trailing * matches a prefix)
 * (gpu.h:38-66 g_gpu_kernel_profile).  Returns dispatches and total ns. */
void gpu_hip_profile_enable(B32 enabled);
void gpu_hip_profile_reset(void);
U64 gpu_hip_profile_query(const char* name, U64* out_dispatches);
/* prefix "tag:" to the profile names of this thread's next dispatches
 * ("" clears), so one kernel used by several operators is told apart */
void gpu_hip_profile_tag(const char* tag);
/* A batch run as a HIP graph (the reference submits each batch as a
 * recorded Vulkan command buffer, gpu_vulkan.c batch_begin/end; graphs
 * have no counterpart there).  The batch's operations are recorded and,
 * at gpu_batch_end, replayed as the calling thread's graph for `key` when
 * they match it (same buffers, sizes, kernels, args and grids; only the
 * bytes written may differ), else captured into a new graph for `key`.
 * Writes and reads go through pinned staging that the graph owns.  Graphs
 * are per host thread, at most 16, least recently used evicted. */
GPU_Batch* gpu_hip_batch_begin_graph(const char* key);
/* graph replays and captures since the process started */
void gpu_hip_graph_stats(U64* replays, U64* captures);
/* pinned host memory (hipHostMalloc): what host-backed columns live in, so
 * their chunk copies are asynchronous DMA (registering existing memory,
 * the reference's host-pointer import, gpu_vulkan.c:2213-2358, is not
 * used: it is refused on some systems) */
void* gpu_hip_host_alloc(U64 bytes);
void gpu_hip_host_free(void* p);
/* called when a device allocation fails even after the pooled buffers were
 * freed: returns the bytes it released, and the allocation is tried once
 * more (g_gpu_evict_hook, query_exec.c:1817-1822) */
void gpu_hip_set_evict_hook(U64 (*hook)(U64 bytes_wanted));
/* pinned host memory from gpu_hip_host_alloc as a buffer the kernels read
 * in place over the host link (the reference's host-pointer import,
 * gpu_vulkan.c:2213-2358, without the copy); releasing it frees nothing */
GPU_Buffer* gpu_hip_buffer_wrap_host(void* host_ptr, U64 size);
/* bytes [offset, offset + size) of a buffer as a buffer of its own, valid
 * while the parent lives; releasing it frees nothing */
GPU_Buffer* gpu_hip_buffer_view(GPU_Buffer* parent, U64 offset, U64 size);
/* fault injection: fail the k-th device allocation from now (gpu.h:68) */
void gpu_hip_fault_alloc_at(U64 k);
/* fault injection: the k-th device-to-host read from now (a synchronous
 * read, or one recorded in a batch, which then fails) reports failure; no
 * counterpart in the reference, whose reads cannot be made to fail */
void gpu_hip_fault_read_at(U64 k);
/* buffer_write / buffer_read that report failure (0, gpu_hip_last_error())
 * instead of dropping it or zero-filling the destination; every operator
 * of this library copies through these */
B32 gpu_hip_buffer_write_checked(GPU_Buffer* buffer, const void* data, U64 size);
B32 gpu_hip_buffer_read_checked(GPU_Buffer* buffer, void* data, U64 size);
const char* gpu_hip_last_error(void);

/* ------------------------------------------------------------------ */
/* Kernel argument block (the launch ABI of every library and run-time */
/* compiled kernel).  The reference binds up to 16 storage buffers by  */
/* index and 8 U64 push constants read as 16 u32 words               */
/* (gpu_vulkan.h:6-9); a HIP kernel takes them by value, with each     */
/* buffer's byte size, so shaders' .length() reads carry over.         */
/* ------------------------------------------------------------------ */
#define QK_MAX_BINDINGS 16
/* 16 words = the reference's 8 x U64 push constants; words 16..31 carry the
 * operator parameters of the aggregate/join kernels (no U64 slot overlaps
 * them) */
#define QK_PUSH_WORDS 32
typedef struct QK_Args {
  U64 ptr[QK_MAX_BINDINGS];    /* device addresses by binding index */
  U64 bytes[QK_MAX_BINDINGS];  /* byte sizes by binding index */
  U32 words[QK_PUSH_WORDS];    /* push constants: U64 slot i = words[2i] | words[2i+1] << 32 */
  U32 global_work_size;        /* threads asked for by the dispatch */
  U32 grid_blocks;             /* workgroups launched */
} QK_Args;

/* ------------------------------------------------------------------ */
/* 2. Operators (query_exec.c hot path)                                 */
/* ------------------------------------------------------------------ */

/* column types: CQL_ColumnType numbering, cql.h:78-95 */
enum {
  CQL_ColumnType_U32 = 1, CQL_ColumnType_U64 = 2, CQL_ColumnType_F32 = 3, CQL_ColumnType_F64 = 4,
  CQL_ColumnType_I32 = 7, CQL_ColumnType_I64 = 8,
};

/* aggregate function codes, query_exec.h:904-910; result types follow
 * qe_agg_func_policy (query_exec.h:917-926): COUNT -> U64, others F64 */
enum { QE_AGG_FUNC_COUNT = 0, QE_AGG_FUNC_SUM = 1, QE_AGG_FUNC_AVG = 2, QE_AGG_FUNC_MIN = 3, QE_AGG_FUNC_MAX = 4 };
#define QE_AGG_MAX_EXPRS 8
#define QE_AGG_MAX_GROUP_COLS 4 /* query_exec.h:898 */

/* WHERE tree as the planner hands it over (IR_Node, flattened in prefix
 * order with child counts).  OPERATOR nodes carry an op; COLUMN nodes a
 * column index of the scanned table; NUMERIC nodes a literal; INLIST holds
 * NUMERIC children (the right side of IN / NOT IN). */
enum { QE_IR_OPERATOR = 1, QE_IR_COLUMN = 2, QE_IR_NUMERIC = 3, QE_IR_INLIST = 4 };
enum {
  QE_IROP_AND = 1, QE_IROP_OR = 2, QE_IROP_EQ = 3, QE_IROP_NE = 4, QE_IROP_LT = 5, QE_IROP_GT = 6,
  QE_IROP_LE = 7, QE_IROP_GE = 8, QE_IROP_IS_NULL = 9, QE_IROP_IS_NOT_NULL = 10, QE_IROP_IN = 11,
  QE_IROP_NOT_IN = 12,
};
typedef struct QE_IRNode {
  U32 type;
  U32 op;
  U32 column;
  U32 child_count;
  F64 value;
} QE_IRNode;

/* scan kernel choice, in the order of qe_scan_filter_selected's chain
 * (query_exec.c:2566-2656) */
enum {
  QE_ScanPath_None = 0,
  QE_ScanPath_GeneratedSimple = 1, /* qk_simple_kernel via hiprtc */
  QE_ScanPath_Generated = 2,       /* qk_scan_kernel via hiprtc */
  QE_ScanPath_Simple = 3,          /* library scan_simple */
  QE_ScanPath_Interpreter = 4,     /* library scan_filter bytecode interpreter */
  QE_ScanPath_ZoneEmpty = 5,       /* every zone pruned, nothing dispatched */
  QE_ScanPath_Constant = 6,        /* predicate folded to a constant */
};

/* per-operator trace (QE_ScanTrace/QE_AggregateTrace, query_exec.h:133-311) */
typedef struct QE_Trace {
  U32 path;              /* QE_ScanPath_* for scans, QE_AggPath_* for aggregates */
  U32 kernel_count;      /* dispatches made */
  U64 rows_before;
  U64 rows_after;
  U64 rows_scanned;      /* rows left after zone-map pruning */
  U64 bytes_read;        /* algorithmic column bytes the kernels read */
  U64 gpu_kernel_ns;     /* summed hipEvent time of the dispatches (profiling on) */
  U64 wall_ns;           /* prepare + execute + fetch */
  U64 compile_ns;        /* hiprtc time spent in this call */
  char kernel[64];       /* name of the main kernel */
} QE_Trace;

enum {
  QE_AggPath_None = 0,
  QE_AggPath_Global = 10,        /* aggregate_global (whole columns) */
  QE_AggPath_FusedGlobal = 11,   /* WHERE + global aggregate in one kernel */
  QE_AggPath_Dense = 12,         /* dense GROUP BY, LDS-private tables */
  QE_AggPath_DenseKeyFold = 13,  /* dense GROUP BY whose args are the key */
  QE_AggPath_FusedDense = 14,    /* WHERE + dense GROUP BY */
  QE_AggPath_JoinDense = 20,     /* join: fact reduced per key value, combined per dim row */
  QE_AggPath_JoinProbeGroup = 21,/* join: hash probe with the dim group code as payload */
  QE_AggPath_JoinProbeRow = 22,  /* join: hash probe, totals per dim row */
  QE_AggPath_JoinPartGroup = 23, /* join: radix-partitioned, LDS probe, group payload */
  QE_AggPath_JoinPartRow = 24,   /* join: radix-partitioned, LDS probe, totals per dim row */
  QE_AggPath_Hash = 30,          /* hash GROUP BY, global open addressing */
};

typedef struct QE_Table QE_Table;

/* settings (settings.c:158-242): key:value switches read at use sites */
void qe_settings_set_u64(const char* key, U64 value);
U64 qe_settings_u64(const char* key, U64 default_value);
void qe_settings_reset(void);

const char* qe_last_error(void);

/* device-resident tables (CQL_Table/CQL_Column, cql.h:206-322, with the
 * columns held whole in HBM; zone maps of CQL_ZONEMAP_CHUNK_ROWS rows,
 * cql.h:70-71, cql.c:3818) */
QE_Table* qe_table_alloc(U64 row_count);
void qe_table_release(QE_Table* table);
U64 qe_table_row_count(QE_Table* table);
/* returns the column index or -1; host_data may be 0 (column left on the
 * device uninitialised, for the generators); null_flags: byte per row,
 * nonzero = NULL (cql.h:283) */
S32 qe_table_add_column(QE_Table* table, const char* name, U32 type, const void* host_data,
                        const U8* null_flags);
GPU_Buffer* qe_table_column_buffer(QE_Table* table, U32 column);
B32 qe_table_column_read(QE_Table* table, U32 column, U64 row0, U64 rows, void* out);
/* recompute zone maps (8192-row min/max) of a column on the device */
B32 qe_table_build_zonemap(QE_Table* table, U32 column);

/* host-backed tables (the reference's streamed cold path, query_exec.c:
 * 1674-2491; columns in host memory, cql.h:283-322).  The rows are copied
 * into pinned host memory; zone maps are built on the device as the rows
 * pass through once.  A table holds either host-backed or device columns.
 * qe_scan_filter and qe_aggregate over such a table stream it in chunks of
 * QE_SCAN_STREAM_CHUNK_ROWS rows (default 16,777,216, query_exec.c:2096-2099)
 * through the resident chunk cache, copies of one chunk overlapping the
 * kernels of the one before.  The other read-only operators (gather,
 * SELECT, joins, batches) read its rows in place over the host link;
 * generators, encoding and device columns are refused. */
S32 qe_table_add_host_column(QE_Table* table, const char* name, U32 type, const void* data);

/* the resident chunk cache (QE_ResidentStats, query_exec.c:1700-1754):
 * budget QE_RESIDENT_BUDGET_PERCENT (70) of HBM, at most
 * QE_RESIDENT_BUDGET_BYTES when set; QE_RESIDENT 0 turns it off */
typedef struct QE_ResidentStats {
  U64 hit_bytes;
  U64 miss_bytes;
  U64 evicted_bytes;
  U64 bytes_in_use;
  U64 entries;
  U64 budget_bytes;
} QE_ResidentStats;
void qe_resident_stats(QE_ResidentStats* out);
/* free every chunk no query holds (a cold start, bench_common.h:516-560) */
void qe_resident_clear(void);

/* the last streamed operator of the calling thread */
typedef struct QE_StreamStats {
  U64 chunks;           /* chunks of the table */
  U64 chunks_skipped;   /* every zone pruned: neither copied nor run */
  U64 chunks_run;
  U64 chunks_resident;  /* every column read from the cache */
  U64 transient_chunks; /* column chunks streamed outside the cache */
  U64 bytes_copied;     /* host to device */
  U64 copy_wait_ns;     /* time the compute side waited for copies */
  U64 wall_ns;
  U64 resident_whole;   /* 1: every column whole in the cache, read in one pass */
} QE_StreamStats;
void qe_stream_last_stats(QE_StreamStats* out);

/* synthetic tables generated in HBM (same definitions as the oracle's) */
B32 qe_gen_events(QE_Table* table, U64 seed, S32 id_col, S32 region_col, S32 amount_col, S32 ts_col);
B32 qe_gen_scale_events(QE_Table* table, S32 id_col, S32 region_col, S32 amount_col, S32 key_col);
B32 qe_gen_fact_fk(QE_Table* table, U32 fk_col, U64 seed, QE_Table* dim, U32 dim_key_col);
B32 qe_gen_dim(QE_Table* table, U64 seed, U32 attr_range, B32 spread_keys, S32 key_col, S32 attr_col);

/* qe_scan_filter_selected (query_exec.c:2495-3273): WHERE over the table;
 * count_only returns just the count; otherwise row ids (U64) are written to
 * out_ids (cap entries), sorted ascending. */
B32 qe_scan_filter(QE_Table* table, const QE_IRNode* where, U32 node_count, B32 count_only,
                   U64* out_ids, U64 cap, U64* out_count, QE_Trace* trace);

/* qe_aggregate_impl (query_exec.c:10258) for the hot-path shapes: optional
 * WHERE, 0..QE_AGG_MAX_GROUP_COLS GROUP BY columns (group_cols, K =
 * group_count; query_exec.c:10280-10299), up to 8 aggregates (arg_cols[e]
 * < 0 = COUNT(*)).  Writes out_group_count groups: out_keys[g*K+k] (key
 * values as F64), out_cells[g*E+e] (U64 bits for COUNT, F64 bits
 * otherwise).  Up to 3 32-bit integer keys whose zone-map ranges multiply
 * to a small range run dense (their digits make one group number, as
 * query_exec.c:9462-9500 does; groups in ascending key order, key 0 most
 * significant); other keys run through the hash table (groups in table-
 * slot order, trace->path QE_AggPath_Hash), as does a WHERE that is not an
 * AND of comparisons.  A SUM over a whole-number column that reaches 2^53
 * fails the query (qe_aggregate_check_sum_exact, query_exec.c:7028-7049;
 * setting QE_WIDE_INT_CHECK, default 1). */
B32 qe_aggregate(QE_Table* table, const QE_IRNode* where, U32 where_count, const S32* group_cols,
                 U32 group_count, const U32* funcs, const S32* arg_cols, U32 expr_count, U64 cap,
                 U64* out_group_count, F64* out_keys, U64* out_cells, QE_Trace* trace);

/* qe_aggregate without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): an
 * ungrouped COUNT/SUM/AVG/MIN/MAX over up to 4 argument columns of a
 * device-resident table, with a WHERE that is an AND of at most 8
 * column-literal comparisons (or none), computed by the same kernels as
 * qe_aggregate, whose result it equals bit for bit.  Its scratch (leaf
 * words, per-workgroup partials, the folded result) is a static device
 * array of QE_NOALLOC_SCRATCH_BYTES that each GPU holds until the process
 * ends; it calls no hipMalloc/hipFree and no allocating library routine.
 * Larger shapes (more leaves or argument columns, other WHERE forms,
 * host-backed tables) are refused with an error.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
#define QE_NOALLOC_SCRATCH_BYTES (256u * 1024u)
B32 qe_aggregate_noalloc(QE_Table* table, const QE_IRNode* where, U32 where_count, const U32* funcs,
                         const S32* arg_cols, U32 expr_count, U64* out_cells, QE_Trace* trace);

/* qe_aggregate_join_pushdown (query_exec.c:8442-8775):
 *   SELECT [dim.g1, ...,] aggs(fact.*) FROM fact JOIN dim ON fact.k = dim.k
 *   [WHERE dim-predicate] [GROUP BY dim.g1, ...]
 * with 0..QE_AGG_MAX_GROUP_COLS U32/I32 dim group columns (K =
 * dim_group_count, query_exec.c:8588-8607; 0 means one group).  Output as
 * qe_aggregate (out_keys[g*K+k]), groups in ascending order of their
 * values, column 0 first.  A fact-side SUM reaching 2^53 fails the query
 * as in qe_aggregate. */
B32 qe_aggregate_join(QE_Table* fact, U32 fact_key_col, QE_Table* dim, U32 dim_key_col,
                      const QE_IRNode* dim_where, U32 dim_where_count, const S32* dim_group_cols,
                      U32 dim_group_count, const U32* funcs, const S32* fact_arg_cols, U32 expr_count, U64 cap,
                      U64* out_group_count, F64* out_keys, U64* out_cells, QE_Trace* trace);

/* Join types of qe_hash_join (the reference's join_type strings "inner",
 * "left", "semi", "anti", query_exec.c:14613-14650). */
enum {
  QE_JoinType_Inner = 0,
  QE_JoinType_Left = 1,
  QE_JoinType_Semi = 2,
  QE_JoinType_Anti = 3,
};

/* qe_hash_join_impl (query_exec.c:14541-15358) for numeric keys: the
 * equi-join probe.key = build.key producing row pairs.  Keys compare as
 * f64 values (-0.0 equal to 0.0, NaN equal to nothing).  probe_rows /
 * build_rows list the rows taking part (NULL = every row of the table;
 * probe positions index probe_rows, or are rows when it is NULL).
 *   inner / left: out holds (probe position, build row) u64 pairs, a left
 *     join's unmatched probe position paired with ~0 (the reference's
 *     BUILD_NULL); every build row with the key fans out.
 *   semi / anti: out holds the probe positions with / without a match, once.
 * Records are sorted by probe position, then by build position.  At most
 * cap records are written; *out_count is the full count.  Returns 0 with
 * qe_last_error() on failure (non-numeric keys, rows out of range, more
 * than 2^31 build rows or 2^32 records). */
U32 qe_hash_join(QE_Table* probe, U32 probe_key_col, const U64* probe_rows, U64 probe_count,
                 QE_Table* build, U32 build_key_col, const U64* build_rows, U64 build_count,
                 U32 join_type, U64* out, U64 cap, U64* out_count, QE_Trace* trace);

/* Narrow column encoding (qe_column_encoding / qe_encode_chunk,
 * query_exec.c:1506-1672): keep a copy of a u32/i32 column as 1-byte or
 * 2-byte codes (value - base, base = the zone-map minimum, when the range is
 * below 256 / 65536), or, with allow_sequence, recognise a sequence (value
 * = first value + row, nothing stored).  scan_simple, the global aggregate
 * and the dense GROUP BY then read the codes while QE_ENCODE is on (the
 * default); other kernels read the full column.  Returns the encoding (0
 * none applies, 1 u8, 2 u16, 3 sequence) or -1 on error (values outside
 * the zone-map range).  Rewriting the column's data drops the encoding. */
S32 qe_table_encode_column(QE_Table* table, U32 column, B32 allow_sequence);

/* Row materialisation (selection_gather.comp modes F32/F64,
 * qe_selection_gather_column): out[c][i] = column cols[c] at row
 * rows[i * stride + offset], at the column's own width (4 or 8 bytes), 0
 * where that id is ~0 (no row, e.g. a left join's unmatched side).  With
 * stride 2 the (probe, build) pairs of qe_hash_join serve as rows
 * directly (offset 0 probe side, 1 build side). */
U32 qe_gather(QE_Table* table, const U32* cols, U32 col_count, const U64* rows, U64 count, U64 stride, U64 offset,
              void* const* out);

/* SELECT cols FROM table WHERE ...: the matching rows' values, in row
 * order, with the row ids kept on the device between the scan and the
 * gather (the reference's device selections, query_exec.h:383-415).  At
 * most cap rows are written per column; *out_count is the full count. */
U32 qe_scan_materialize(QE_Table* table, const QE_IRNode* where, U32 where_count, const U32* cols, U32 col_count,
                        U64 cap, U64* out_count, void* const* out);

/* Batched shared scan.  No counterpart in the reference, whose fused
 * kernels serve one query each (query_kernel.c:1433-1600): many aggregate
 * queries over the same table run as one pass over the union of their
 * columns.  Each query is what qe_aggregate takes, and each result is what
 * qe_aggregate returns for it alone (f64 sums within 1e-12 relative: the
 * summation order differs).  Queries fuse when the WHERE is an AND of at
 * most 8 comparisons, there is at most one GROUP BY key and it is a 32-bit
 * integer column whose zone-map range is at most QE_BATCH_MAX_GROUPS
 * (1024), and they take at most 2 distinct argument columns; other
 * queries (several keys among them) run through qe_aggregate inside the
 * same call.  `ok` is set per query. */
typedef struct QE_BatchQuery {
  const QE_IRNode* where;
  U32 where_count;
  U32 group_count;                       /* GROUP BY columns, 0..QE_AGG_MAX_GROUP_COLS */
  S32 group_cols[QE_AGG_MAX_GROUP_COLS];
  const U32* funcs;
  const S32* arg_cols;
  U32 expr_count;
  U64 cap;
  U64* out_group_count;
  F64* out_keys;  /* cap x group_count key values (may be 0) */
  U64* out_cells; /* cap x expr_count cells */
  B32 ok;
  U32 fused; /* out: 1 when the query ran in a shared scan */
} QE_BatchQuery;

typedef struct QE_BatchTrace {
  U32 queries;
  U32 fused;
  U32 fallback;
  U32 launches;     /* shared-scan launches made */
  U32 items;        /* rows per thread per tile of the shared scan */
  U32 blocks_per_cu;
  U64 rows_scanned; /* rows read per launch */
  U64 bytes_read;   /* column bytes the launches read (union per launch) */
  U64 gpu_kernel_ns;
  U64 wall_ns;
  U32 tuned_wg;      /* configuration used: block size asked for (0 = first that fits) */
  U32 tuned_buffers; /* stage buffers asked for (0 one, 1 two, 2 two when they fit) */
} QE_BatchTrace;

/* launch configuration of the shared scan: block size (256/512/1024, 0 =
 * the largest whose stage fits), rows per thread staged (at most), stage
 * buffers (0 one, 1 two, 2 two when they fit), blocks per CU, and the LDS
 * bytes a launch's stage and accumulators may take (0 = a CU's LDS) */
typedef struct QE_BatchConfig {
  U32 wg, items, buffers, blocks_per_cu;
  U64 lds_budget;
} QE_BatchConfig;

/* With setting QE_BATCH_AUTOTUNE 1 the configuration comes from the
 * autotuner (tuned on the first batch of each workload: table size, width
 * of the queries' columns, query count); otherwise from the QE_BATCH_*
 * settings. */
B32 qe_aggregate_batch(QE_Table* table, QE_BatchQuery* queries, U32 count, QE_BatchTrace* trace);
/* Autotuner: times the candidate configurations on this batch, records the
 * fastest for its workload (in memory and in $CQLHIP_TUNE_FILE, default
 * ~/.cache/cqlhip/batch_tune.txt) and returns it with its time. */
B32 qe_aggregate_batch_tune(QE_Table* table, QE_BatchQuery* queries, U32 count, QE_BatchConfig* best,
                            F64* best_ms);

/* Served request loop.  Replaces, for aggregate queries, the one-query-
 * at-a-time execution of the reference's Postgres-wire server (pg_server.c:
 * 200-250, :643-760, serialised by g_query_exec_mutex, application.h:34):
 * queued requests are batched into shared scans by `workers` threads (one
 * batch in flight each), up to `max_batch` queries per batch, admitting
 * requests while the device memory they need stays within `hbm_budget`;
 * requests arriving while batches run join the next batch.  max_batch 1
 * runs each request alone through qe_aggregate (the plain path).  A
 * request's QE_BatchQuery and the arrays it points to stay owned by the
 * caller and must live until qe_server_wait returns 1 for its ticket;
 * results are what qe_aggregate_batch returns for them. */
typedef struct QE_Server QE_Server;

typedef struct QE_ServerConfig {
  U32 workers;    /* default 2 */
  U32 max_batch;  /* default 64 */
  U64 hbm_budget; /* default 1 GiB */
} QE_ServerConfig;

typedef struct QE_ServerStats {
  U64 submitted, completed, failed;
  U64 batches, batched_queries;
  U64 p50_ns, p99_ns, max_ns, mean_ns; /* submit to completion */
  U64 elapsed_ns;                      /* first submit to last completion */
  U64 inflight_peak_bytes, queue_peak;
  U64 joins, scans; /* completed requests of those kinds */
} QE_ServerStats;

QE_Server* qe_server_start(const QE_ServerConfig* config);
/* returns a ticket, 0 on error */
U64 qe_server_submit(QE_Server* server, QE_Table* table, QE_BatchQuery* query);
/* 1 when the request is done (query->ok says whether it succeeded; the
 * ticket is then released), 0 on timeout */
B32 qe_server_wait(QE_Server* server, U64 ticket, U64 timeout_ns);
void qe_server_stats(QE_Server* server, QE_ServerStats* out);
void qe_server_stats_reset(QE_Server* server);
/* runs what is queued, then stops the workers and frees the server */
void qe_server_stop(QE_Server* server);
/* load driver: Poisson arrivals at rate_qps for duration_ns (seeded),
 * request k a copy of pool[k % pool_count] with output buffers of its own
 * (pool entries' out_* fields are ignored; cap sizes the buffers);
 * returns once every request has finished, with the server's stats over
 * the run (they are reset at its start). */
B32 qe_server_drive(QE_Server* server, QE_Table* table, const QE_BatchQuery* pool, U32 pool_count, F64 rate_qps,
                    U64 duration_ns, U64 seed, QE_ServerStats* out);
/* the same load over several tables: request k runs pool[k % pool_count]
 * on tables[k % pool_count] (a mix of tables and query shapes) */
B32 qe_server_drive_mix(QE_Server* server, QE_Table* const* tables, const QE_BatchQuery* pool, U32 pool_count,
                        F64 rate_qps, U64 duration_ns, U64 seed, QE_ServerStats* out);

/* Served joins and scans.  Besides aggregates, the reference's server runs
 * every other query its clients send through the same one-at-a-time
 * executor (pg_server.c:643-760): here a pushed-down join aggregate
 * (qe_aggregate_join's arguments; the submit call's table is the fact
 * table) and a row-id scan or count (qe_scan_filter's, the point lookup
 * among them).  Each runs alone on a worker while the other workers run
 * aggregate batches, and is admitted by its device footprint like them
 * (a join's partition buffers: fact rows x 4 B x (1 + arguments) plus its
 * dimension's table; a scan's id buffer).  The struct and the arrays it
 * points to stay owned by the caller until qe_server_wait returns 1. */
typedef struct QE_JoinQuery {
  QE_Table* dim;
  U32 fact_key_col, dim_key_col;
  const QE_IRNode* dim_where;
  U32 dim_where_count;
  U32 group_count; /* dim GROUP BY columns, 0..QE_AGG_MAX_GROUP_COLS */
  S32 group_cols[QE_AGG_MAX_GROUP_COLS];
  const U32* funcs;
  const S32* arg_cols; /* fact columns */
  U32 expr_count;
  U64 cap;
  U64* out_group_count;
  F64* out_keys; /* cap x group_count key values (may be 0) */
  U64* out_cells; /* cap x expr_count cells */
  B32 ok;
} QE_JoinQuery;

typedef struct QE_ScanQuery {
  const QE_IRNode* where;
  U32 where_count;
  B32 count_only;
  U64 cap;       /* out_ids capacity (ignored when count_only) */
  U64* out_ids;  /* matching row ids, ascending */
  U64* out_count;
  B32 ok;
} QE_ScanQuery;

/* Many joins over one fact table at once (no counterpart in the
 * reference, which runs each join alone: query_exec.c:8442-8775).
 * Identical joins run once (QE_BATCH_DEDUPE); joins that need the same
 * fact side share it: on the partitioned path the fact rows in partition
 * order (same dim size, fact key and argument columns; each join builds
 * its own table and probes), on the dense path the per-key totals (same
 * fact key and arguments; each join combines them with its own dim rows).
 * Every query's `ok` says whether it ran; the call returns 0 only for bad
 * arguments.  Results are what qe_aggregate_join returns for each. */
typedef struct QE_JoinBatchTrace {
  U32 queries;     /* joins in the batch */
  U32 runs;        /* joins run (identical ones folded) */
  U32 fact_passes; /* fact sides computed; the other runs reused one */
  U64 wall_ns;
} QE_JoinBatchTrace;
B32 qe_aggregate_join_batch(QE_Table* fact, QE_JoinQuery* queries, U32 count, QE_JoinBatchTrace* trace);

U64 qe_server_submit_join(QE_Server* server, QE_Table* fact, QE_JoinQuery* query);
U64 qe_server_submit_scan(QE_Server* server, QE_Table* table, QE_ScanQuery* query);

enum { QE_REQUEST_AGGREGATE = 0, QE_REQUEST_JOIN = 1, QE_REQUEST_SCAN = 2 };

typedef struct QE_ServeRequest {
  U32 kind;          /* QE_REQUEST_* */
  QE_Table* table;   /* the table scanned (a join's fact table) */
  const void* query; /* QE_BatchQuery, QE_JoinQuery or QE_ScanQuery */
} QE_ServeRequest;

/* the load driver over requests of every kind: request k is a copy of
 * pool[k % pool_count] with output buffers of its own, sized by its cap */
B32 qe_server_drive_requests(QE_Server* server, const QE_ServeRequest* pool, U32 pool_count, F64 rate_qps,
                             U64 duration_ns, U64 seed, QE_ServerStats* out);

/* bandwidth probe (bandwidth_read.comp): reads `bytes` of a buffer with
 * 16-byte loads; returns a sink value.  For measuring the HBM ceiling. */
B32 qe_bandwidth_read(GPU_Buffer* buffer, U64 bytes, U32 blocks, U64* out_sink, QE_Trace* trace);
/* the copy counterpart: `bytes` of src written to dst (the read + write
 * ceiling the partition scatter is measured against) */
B32 qe_bandwidth_copy(GPU_Buffer* src, GPU_Buffer* dst, U64 bytes, U32 blocks, QE_Trace* trace);

#ifdef __cplusplus
}
#endif
#endif
