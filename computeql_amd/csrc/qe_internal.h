/*
 * qe_internal.h - shared declarations of the host operator layer (C99).
 * The operators drive the GPU only through the GPU_Backend table, the way
 * query_exec.c drives it through the gpu_* wrappers (gpu.h:95-159).
 */
#ifndef QE_INTERNAL_H
#define QE_INTERNAL_H

#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include "../../include/cqlhip.h"
#include "qk_abi.h"

#define QE_MAX_COLUMNS 64
#define QE_ZONE_SHIFT 13 /* CQL_ZONEMAP_CHUNK_ROWS 8192, cql.h:70-71 */
#define QE_SCAN_MAX_LEAVES 8
#define QE_MAX_COLUMN_BINDINGS 10 /* query_exec.h:30 */
#define QE_WG 256

typedef struct QE_Column {
  char name[64];
  U32 type;
  U32 width; /* bytes per row */
  GPU_Buffer* buffer;
  B32 has_nulls;
  /* with NULLs: byte offset in `buffer` of the null bitmap (bit r of word
   * r / 32 set = row r is NULL), placed after the rows and their slack */
  U64 null_offset;
  /* zone maps: min/max per 8192 rows, and of the whole column */
  U64 zone_count;
  F64* zone_min;
  F64* zone_max;
  B32 zonemap_valid;
  F64 min_value;
  F64 max_value;
  /* groups of 64 zones: the min and max of the zones' minima and of their
   * maxima (4 values a group), which decide most groups without looking
   * at their zones (qe_zonemap_prune); 0 when absent */
  F64* super_zone;
  U64 super_count;
  /* narrow encoding (qe_table_encode_column): enc 0 none, 1 u8 codes, 2 u16
   * codes (in enc_buffer), 3 sequence (value = enc_base + row); dropped
   * whenever the column's data is rewritten */
  U32 enc;
  U32 enc_base;
  GPU_Buffer* enc_buffer;
  /* host-backed column (qe_table_add_host_column): the rows live in pinned
   * host memory and reach the device chunk by chunk through the resident
   * chunk cache, under `uid` (buffer is 0) */
  void* host_data;
  U64 uid;
} QE_Column;

struct QE_Table {
  U64 row_count;
  U32 column_count;
  B32 host_backed; /* every column is host-backed; operators stream it */
  QE_Column columns[QE_MAX_COLUMNS];
};

U32 qe_type_width(U32 type); /* bytes per row of a column type, 0 if unsupported */

/* operators that read column buffers: a host-backed table serves them
 * through its zero-copy buffers (0 when the device cannot map them) */
B32 qe_require_device(const QE_Table* t, const char* op);
/* operators that write columns or add device state refuse host-backed tables */
B32 qe_refuse_host(const QE_Table* t, const char* op);
/* a query touching a column with NULLs is refused (column < 0: none) */
B32 qe_require_non_null(const QE_Table* t, S32 column, const char* op);
B32 qe_where_non_null(const QE_Table* t, const QE_IRNode* where, U32 count, const char* op);

/* ---- streamed execution over host-backed tables (qe_stream.c) ---- */
B32 qe_stream_scan(QE_Table* table, const QE_IRNode* where, U32 node_count, B32 count_only, U64* out_ids, U64 cap,
                   U64* out_count, QE_Trace* trace);
B32 qe_stream_aggregate(QE_Table* table, const QE_IRNode* where, U32 where_count, const S32* group_cols,
                        U32 group_count, const U32* funcs, const S32* arg_cols, U32 expr_count, U64 cap,
                        U64* out_group_count, F64* out_keys, U64* out_cells, QE_Trace* trace);
/* drop the cached chunks of a column (its table is being released) */
void qe_resident_drop_column(U64 uid);

/* ---- gpu_* wrappers over the active backend (gpu_dispatch.c:153-603) ---- */
GPU_Backend* qe_gpu(void);
#define gpu_buffer_alloc(s, f, d) (qe_gpu()->buffer_alloc((s), (f), (d)))
#define gpu_buffer_release(b) (qe_gpu()->buffer_release((b)))
#define gpu_kernel_alloc(n) (qe_gpu()->kernel_alloc((n)))
#define gpu_kernel_alloc_from_spirv(n, c) (qe_gpu()->kernel_alloc_from_spirv((n), (c)))
#define gpu_kernel_set_arg_buffer(k, i, b) (qe_gpu()->kernel_set_arg_buffer((k), (i), (b)))
#define gpu_kernel_set_arg_u64(k, i, v) (qe_gpu()->kernel_set_arg_u64((k), (i), (v)))
#define gpu_batch_begin(u, d) (qe_gpu()->batch_begin((u), (d)))
#define gpu_batch_buffer_write(b, buf, d, s) (qe_gpu()->batch_buffer_write((b), (buf), (d), (s)))
#define gpu_batch_buffer_zero(b, buf, s) (qe_gpu()->batch_buffer_zero((b), (buf), (s)))
#define gpu_batch_buffer_fill(b, buf, s, v) (qe_gpu()->batch_buffer_fill((b), (buf), (s), (v)))
#define gpu_batch_kernel_execute(b, k, g, l) (qe_gpu()->batch_kernel_execute((b), (k), (g), (l)))
#define gpu_batch_buffer_read(b, buf, d, s) (qe_gpu()->batch_buffer_read((b), (buf), (d), (s)))
#define gpu_batch_end(b) (qe_gpu()->batch_end((b)))

/* synchronous copies that fail the operator (0, qe_last_error()) when the
 * copy fails; the backend's buffer_read zero-fills instead */
B32 qe_buffer_read_checked(GPU_Buffer* b, void* out, U64 size);
B32 qe_buffer_write_checked(GPU_Buffer* b, const void* data, U64 size);

/* the no-allocation scratch of qe_aggregate_noalloc (gpu_hip.cpp): its
 * device address on the current device (0 when not set up), and a run of
 * library kernels over it that allocates nothing */
U64 gpu_hip_noalloc_scratch(void);
B32 gpu_hip_noalloc_run(const void* upload, U64 upload_bytes, U64 upload_offset, const char* const* kernels,
                        const QK_Args* args, const U32* blocks, U32 steps, void* download, U64 download_bytes,
                        U64 download_offset);

String8 qe_str8(const char* s);
GPU_Kernel* qe_kernel(const char* name);
/* set push words directly (words[i]) */
void qe_kernel_set_word(GPU_Kernel* k, U32 word, U32 value);
/* pooled scratch buffer of at least `bytes`, by name (per thread) */
GPU_Buffer* qe_scratch(const char* name, U64 bytes);
void qe_scratch_release(GPU_Buffer* b);

/* ---- errors / timing ---- */
void qe_set_error(const char* fmt, ...);
U64 qe_now_ns(void);
void qe_trace_begin(QE_Trace* t, U64* t0, U64* prof0);
void qe_trace_end(QE_Trace* t, U64 t0, U64 prof0);

/* ---- grid sizing ---- */
U32 qe_stream_blocks(U64 quads);      /* blocks for a streaming pass over `quads` quads */
U32 qe_setting_blocks_per_cu(void);

/* ---- WHERE compilation (query_exec.c:203-545, :1189-1456) ---- */
typedef struct QE_ColumnSlots {
  U32 count;
  U32 column[QE_MAX_COLUMN_BINDINGS]; /* table column of slot i */
} QE_ColumnSlots;

U32 qe_slot_of(QE_ColumnSlots* slots, U32 column); /* ~0u when out of slots */

/* integer-leaf flags of word 1 (query_exec.c:1372-1374) */
#define QE_SCAN_INT_FLAG 0x10000u
#define QE_SCAN_INT_NEVER 0x20000u
#define QE_SCAN_INT_BIAS 0x40000u
/* leaf ops past the five comparisons: IS NULL (5) and IS NOT NULL (6),
 * flagged in word 1 so no value is compared */
#define QE_LEAF_IS_NULL 5u
#define QE_LEAF_IS_NOT_NULL 6u
#define QE_SCAN_NULL_TEST 0x80000u
/* op 7: the value is one of the integers lo + i for the bits i of a 64-bit
 * set (an IN list over a 32-bit integer column): word 2 holds bits 0-31,
 * word 4 bits 32-63, word 3 lo (as for an integer interval) */
#define QE_LEAF_IN_SET 7u
#define QE_SCAN_INT_SET 0x200000u
/* the leaf passes where its comparison fails (!=, NOT IN); a NULL row
 * still fails it */
#define QE_SCAN_NEGATE 0x100000u

/* the encoding kernels are given for a column (0 unless it is encoded and
 * QE_ENCODE is on), and the buffer they read for it */
U32 qe_col_enc(const QE_Table* t, U32 col);
GPU_Buffer* qe_col_kbuf(const QE_Table* t, U32 col);
/* rewrite an integer leaf (qe_simple_leaf_words) to compare codes of an
 * encoded column: lo -= base (and the i32 bias), width in bits 8-15 of
 * word 1 (the reference's leaf layout, query_exec.c:1442-1456) */
void qe_leaf_encode(U32* words, U32 enc, U32 base);

typedef struct QE_Bytecode {
  U32 words[4096];
  U32 word_count;
  F64 consts[256];
  U32 const_count;
  QE_ColumnSlots slots;
  B32 overflow;
  char error[128];
} QE_Bytecode;

B32 qe_bytecode_program_build(QE_Table* table, const QE_IRNode* nodes, U32 count, QE_Bytecode* out);
U32 qe_ir_subtree_end(const QE_IRNode* nodes, U32 i);

typedef struct QE_Leaf {
  U32 column;
  U32 op; /* 0 EQ, 1 LT, 2 LE, 3 GT, 4 GE (query_exec.c QE_CPU_OP_*), 5 IS
             NULL, 6 IS NOT NULL, 7 IN a set (target = its lowest value) */
  F64 target;
  B32 negate; /* != and NOT IN: the comparison's complement, NULL failing */
  U64 set;    /* op 7: bit i = target + i is in the set */
} QE_Leaf;

/* A WHERE of AND, OR, comparisons, IN / NOT IN and IS [NOT] NULL as a
 * disjunction of conjunctions over at most QE_SCAN_MAX_LEAVES distinct
 * leaves: the row passes when every leaf of some term passes.  A leaf over
 * a NULL fails, and AND / OR are two-valued, as the reference's row
 * evaluation has it (qe_row_condition_eval, query_exec.c:13873-14043), so
 * NULLs are exact under OR as well.  term_count 0 = the AND of every
 * leaf; a WHERE that is never true has one leaf that never passes. */
#define QE_WHERE_MAX_TERMS 16
/* device words: 6 per leaf, then the terms' leaf bits, a byte each */
#define QE_WHERE_WORDS (6 * QE_SCAN_MAX_LEAVES + QE_WHERE_MAX_TERMS / 4)
typedef struct QE_Where {
  QE_Leaf leaf[QE_SCAN_MAX_LEAVES];
  U32 leaf_count;
  U32 term_count;
  U8 term[QE_WHERE_MAX_TERMS];
} QE_Where;
/* 0 when the tree has another shape or needs more leaves or terms */
B32 qe_where_build(const QE_Table* t, const QE_IRNode* nodes, U32 count, QE_Where* out);
/* the words the leaf kernels take for `w`, leaf l over column slot
 * slot[l], with NULL bitmaps and column encodings applied when asked;
 * returns the bytes to upload and sets *packed to the leaf-count word the
 * kernels take (leaves | terms << 8) */
U32 qe_where_words(const QE_Table* t, const QE_Where* w, const U32* slot, B32 nulls, B32 enc, U32* words, U32* packed);

/* top-level AND of `column op literal` comparisons; 0 if the tree is not */
B32 qe_collect_and_leaves(const QE_IRNode* nodes, U32 count, QE_Leaf* leaves, U32* leaf_count, U32 max_leaves);
/* conjuncts usable for zone-map pruning (any top-level AND leaf) */
U32 qe_collect_prune_leaves(const QE_IRNode* nodes, U32 count, QE_Leaf* leaves, U32 max_leaves);
void qe_simple_leaf_words(U32* words, U32 slot, U32 type, U32 op, F64 target);
/* word 5 of a leaf over a column with NULLs: its null bitmap's offset in
 * 16-byte units (0 = no NULLs); the leaf kernels then count a NULL row as
 * failing the comparison (query_exec.c:13971-14043) */
void qe_leaf_nulls(U32* words, const QE_Table* t, U32 column);
B32 qe_scan_int_interval(U32 op, F64 target, F64 code_offset, F64 max_code, U32* out_lo, U32* out_span);

/* zone-map pruning: dispatch range [row_begin,row_end) and a skip bitmap
 * (one bit per zone over the whole table, set = skipped); returns the
 * surviving zone count */
typedef struct QE_Prune {
  U64 row_begin, row_end;
  U64 zones_total, zones_left;
  U32* skip_bits; /* malloc'd, zones_total bits, or 0 when nothing skipped */
  U64 skip_words;
} QE_Prune;
void qe_zonemap_prune(QE_Table* table, const QE_Leaf* leaves, U32 leaf_count, QE_Prune* out);
#define QE_SUPER_SHIFT 6 /* zones per zone group: 64 */
/* (re)build a column's zone groups from its zone maps */
void qe_zonemap_build_super(QE_Column* col);
B32 qe_scan_filter_device(QE_Table* table, const QE_IRNode* where, U32 node_count, GPU_Buffer** out_ids, U64* out_count);
void qe_prune_free(QE_Prune* p);

/* ---- generated kernels (query_kernel.c) ---- */
/* the WHERE tree as one HIP predicate kernel; returns 0 (and sets the error)
 * when hiprtc is unavailable or compilation fails */
GPU_Kernel* qk_scan_kernel(QE_Table* table, const QE_IRNode* nodes, U32 count, QE_ColumnSlots* slots,
                           F64* consts, U32* const_count, U64* compile_ns);
/* QE_ROUTE_CONJUNCTIONS: 1 when this WHERE shape's first run goes to the
 * library kernels (its kernel is generated when the shape comes back) */
B32 qk_route_defer(const QE_Bytecode* p);
GPU_Kernel* qk_simple_kernel(QE_Table* table, const QE_Leaf* leaves, U32 leaf_count, QE_ColumnSlots* slots,
                             U64* compile_ns);

/* ---- aggregate result typing (query_exec.h:898-926, query_exec.c:7053) ---- */
typedef struct QE_GroupTotals {
  U64 count;
  S64 isum[QE_AGG_MAX_EXPRS]; /* exact integer sums per distinct arg */
  F64 fsum[QE_AGG_MAX_EXPRS];
  F64 mn[QE_AGG_MAX_EXPRS];
  F64 mx[QE_AGG_MAX_EXPRS];
} QE_GroupTotals;

/* result typing: COUNT -> U64, SUM/AVG/MIN/MAX -> F64 bits
 * (qe_agg_func_policy, query_exec.h:917-926; qe_aggregate_finish,
 * query_exec.c:7269-7293) */
static inline U64 qe_agg_cell(U32 func, U64 count, F64 sum, F64 mn, F64 mx)
{
  F64 v;
  switch (func) {
    case QE_AGG_FUNC_COUNT: return count;
    case QE_AGG_FUNC_SUM: v = sum; break;
    case QE_AGG_FUNC_AVG: v = count ? sum / (F64)count : 0.0; break;
    case QE_AGG_FUNC_MIN: v = mn; break;
    default: v = mx; break;
  }
  U64 bits;
  memcpy(&bits, &v, 8);
  return bits;
}

/* the cell of an aggregate over a 32-bit integer column, its sum exact;
 * with `raw` (qe_t_raw_int_sums of the calling thread, read once per
 * operator: set while streamed aggregation merges chunk results) a SUM
 * cell holds the S64 sum, not its F64 */
static inline U64 qe_agg_cell_int(U32 func, U64 count, S64 isum, F64 mn, F64 mx, B32 raw)
{
  if (raw && func == QE_AGG_FUNC_SUM) return (U64)isum;
  return qe_agg_cell(func, count, (F64)isum, mn, mx);
}
extern __thread B32 qe_t_raw_int_sums;

/* qe_aggregate_check_sum_exact (query_exec.c:7028-7049, from
 * qe_aggregate_build_output :7083): a SUM over a whole-number column
 * (U32/I32/U64/I64) whose result reaches 2^53 in magnitude may have lost
 * low digits as a double, so the query fails (setting QE_WIDE_INT_CHECK,
 * default 1).  cells: rows x expr_count, arg_cols index `t`'s columns. */
B32 qe_check_sum_exact(const QE_Table* t, const U32* funcs, const S32* arg_cols, U32 expr_count, const U64* cells,
                       U64 rows);

/* an aggregate's expressions resolved to distinct argument columns */
#define QE_AGG_MAX_ARGS 4
typedef struct QE_AggPlan {
  U32 expr_count;
  U32 funcs[QE_AGG_MAX_EXPRS];
  S32 expr_arg[QE_AGG_MAX_EXPRS]; /* index into args, -1 for COUNT(*) */
  U32 arg_count;
  U32 arg_col[QE_AGG_MAX_ARGS];
  B32 need_sum, need_minmax;
} QE_AggPlan;
B32 qe_agg_plan(QE_Table* t, const U32* funcs, const S32* arg_cols, U32 expr_count, QE_AggPlan* p);

/* GROUP BY columns with what their zone maps say: for a 32-bit integer
 * key, the code of its minimum (I32 biased by 2^31) and its span; range 0
 * when the key is another type or has no zone map */
typedef struct QE_GroupKeys {
  U32 count;
  U32 col[QE_AGG_MAX_GROUP_COLS];
  U32 type[QE_AGG_MAX_GROUP_COLS];
  U32 min_code[QE_AGG_MAX_GROUP_COLS];
  U64 range[QE_AGG_MAX_GROUP_COLS];
} QE_GroupKeys;
B32 qe_group_keys(const QE_Table* t, const S32* cols, U32 count, QE_GroupKeys* out);
/* the product of the key ranges (0 when a key has none; saturates at 2^64-1) */
U64 qe_group_keys_span(const QE_GroupKeys* k);
/* the value of key k whose digit (code - minimum) is d, as F64 */
F64 qe_group_key_value(const QE_GroupKeys* k, U32 i, U64 d);

/* per-group device totals of aggregate_dense / join_probe: count u64[R];
 * sum u64[E][R]; min u32[E][R]; max u32[E][R] */
typedef struct QE_DenseTotals {
  U32 range, args;
  B32 wide; /* arguments wider than 32-bit integers: sums f64 bits for those, min/max as f64 order keys */
  U64* count;
  U64* sum;
  U32* mn;
  U32* mx;
  U64* mn64;
  U64* mx64;
  void* block;
} QE_DenseTotals;

/* dense GROUP BY over 1..3 32-bit integer keys whose ranges multiply to
 * at most the LDS table's range: group = the keys' digits (key 0 most
 * significant) */
#define QE_DENSE_MAX_KEYS 3
B32 qe_dense_run(QE_Table* t, const QE_GroupKeys* keys, const U32* arg_cols, U32 arg_count, B32 need_sum,
                 B32 need_minmax, const QE_Where* where, const QE_Prune* prune, QE_DenseTotals* out,
                 QE_Trace* trace);
void qe_dense_totals_free(QE_DenseTotals* d);
U64 qe_dense_lds(U32 range, U32 args, B32 need_sum, B32 need_minmax, B32 narrow);
B32 qe_dense_narrow_ok(const QE_Table* t, const U32* arg_cols, U32 args, B32 need_sum, B32 need_minmax, U32 range);
U64 qe_dense_totals_bytes(U32 range, U32 args);

#endif
