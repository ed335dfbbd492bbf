/*
 * oracle.h - CPU restatement of ComputeQL's CPU operator path for the
 * scan+filter -> group-by/aggregate -> join hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker (or as the timed CPU
 * baseline), never as the thing measured or shipped.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the ComputeQL source tree).  Pinning: see the "pinned by" note of each
 * function; every function here is checked by tests/test_oracle.py against
 * closed forms (bench_select_scale.c:98-168), brute force, or invariants.
 */
#ifndef CQL_ORACLE_H
#define CQL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint8_t  orc_u8;
typedef uint32_t orc_u32;
typedef int32_t  orc_s32;
typedef uint64_t orc_u64;
typedef int64_t  orc_s64;
typedef double   orc_f64;

/* column types: same numbering as CQL_ColumnType (cql.h:78-95) */
enum {
  ORC_TYPE_U32 = 1, ORC_TYPE_U64 = 2, ORC_TYPE_F32 = 3, ORC_TYPE_F64 = 4,
  ORC_TYPE_I32 = 7, ORC_TYPE_I64 = 8,
};

/* aggregate func codes: query_exec.h:904-910 */
enum { ORC_AGG_COUNT = 0, ORC_AGG_SUM = 1, ORC_AGG_AVG = 2, ORC_AGG_MIN = 3, ORC_AGG_MAX = 4 };

/* IR node, prefix order (the planner's WHERE tree flattened).  Same layout as
 * QE_IRNode in include/cqlhip.h. */
enum { ORC_IR_OPERATOR = 1, ORC_IR_COLUMN = 2, ORC_IR_NUMERIC = 3, ORC_IR_INLIST = 4 };
enum {
  ORC_OP_AND = 1, ORC_OP_OR = 2, ORC_OP_EQ = 3, ORC_OP_NE = 4, ORC_OP_LT = 5, ORC_OP_GT = 6,
  ORC_OP_LE = 7, ORC_OP_GE = 8, ORC_OP_IS_NULL = 9, ORC_OP_IS_NOT_NULL = 10, ORC_OP_IN = 11,
  ORC_OP_NOT_IN = 12,
};
typedef struct orc_ir_node {
  orc_u32 type;
  orc_u32 op;
  orc_u32 column;
  orc_u32 child_count;
  orc_f64 value;
} orc_ir_node;

/* a column as the CPU path sees it: typed host array + optional byte-per-row
 * NULL flags (cql.h:283, nonzero = NULL) */
typedef struct orc_column {
  orc_u32 type;
  orc_u32 pad;
  const void* data;
  const orc_u8* null_flags;
} orc_column;

/* ---------------- generators ---------------- */
/* splitmix64 step, bench_common.h:30-37; the k-th output of a stream seeded
 * with s is mix(s + (k+1)*0x9E3779B97F4A7C15). */
orc_u64 orc_splitmix64_at(orc_u64 seed, orc_u64 k);

/* seeded 4 x u32 table: id = row, region = d(3r)%50, amount = d(3r+1)%10000,
 * ts = d(3r+2)>>32, d(k) = orc_splitmix64_at(seed, k).  Rows [row0,row0+n). */
void orc_gen_events(orc_u64 seed, orc_u64 row0, orc_u64 n, orc_u32* id, orc_u32* region,
                    orc_u32* amount, orc_u32* ts, int threads);
/* fact foreign key: fk[r] = dim_keys[d'(r) % dim_rows], d' seeded with seed ^ FK salt */
void orc_gen_fact_fk(orc_u64 seed, orc_u64 row0, orc_u64 n, const orc_u32* dim_keys,
                     orc_u64 dim_rows, orc_u32* fk, int threads);
/* dimension: keys = affine bijection of a Fisher-Yates permutation of 0..n-1,
 * attr = d''(j) % attr_range */
void orc_gen_dim(orc_u64 seed, orc_u64 n, orc_u32 attr_range, int spread_keys,
                 orc_u32* keys, orc_u32* attr);
/* closed-form table of bench_select_scale.c:39-45 / :172-197 */
void orc_gen_scale_events(orc_u64 row0, orc_u64 n, orc_u32* id, orc_u32* region,
                          orc_u32* amount, orc_u32* needle_key);

/* closed-form answers, bench_select_scale.c:98-168 */
orc_f64 orc_ss_sum_amount(orc_u64 rows);
orc_u64 orc_ss_count_amount_above(orc_u64 rows, orc_u64 threshold);
orc_u64 orc_ss_count_region(orc_u64 rows, orc_u64 region);
orc_f64 orc_ss_region_amount_sum(orc_u64 rows, orc_u64 region);
orc_u64 orc_ss_amount_rows(orc_u64 rows, orc_u64 amount);
orc_f64 orc_ss_weight_sum(orc_u64 count);
orc_u32 orc_ss_key_of_row(orc_u64 row);

/* ---------------- scan / filter ---------------- */
/* qe_row_condition_eval, query_exec.c:13874-14043 (numeric subset): f64
 * comparisons, NULL operand => false, AND/OR short-circuit, IN list. */
int orc_row_condition_eval(const orc_ir_node* nodes, orc_u32 node_count,
                           const orc_column* cols, orc_u64 row);
/* qe_cpu_scan_filter, query_exec.c:14064-14135: matching row ids in row
 * order.  Returns the count; writes ids when out != 0 (cap entries max). */
orc_u64 orc_cpu_scan_filter(const orc_ir_node* nodes, orc_u32 node_count, const orc_column* cols,
                            orc_u64 rows, orc_u64* out, orc_u64 cap, int threads);

/* qe_scan_int_interval, query_exec.c:1378-1404: op 0 EQ,1 LT,2 LE,3 GT,4 GE */
int orc_scan_int_interval(orc_u32 op, orc_f64 target, orc_f64 code_offset, orc_f64 max_code,
                          orc_u32* out_lo, orc_u32* out_span);

/* ---------------- aggregates ---------------- */
/* qe_aggregate_global_chunked, query_exec.c:9114-9228 with the
 * QE_CHUNK_REDUCE_LOOP of :7472-7482: 64M-row chunks, tasks = max(1,
 * min(workers, rows/65536)) contiguous ranges (tp_divide_work,
 * thread_pool.c:236-249), per-task f64 partials combined in task order.
 * out_sum/min/max per column. */
void orc_aggregate_global_chunked(const orc_column* cols, orc_u32 col_count, orc_u64 rows,
                                  int workers, orc_f64* out_sum, orc_f64* out_min, orc_f64* out_max);

/* qe_aggregate_cpu_reduce, query_exec.c:7297-7399 (+ hash :7201-7234,
 * equality :7237-7266, finish :7269-7293): serial, groups in first-
 * appearance order, accumulators start at min=1e300/max=-1e300.
 * keys: key_count f64 arrays; args: expr_count f64 arrays (0 for COUNT).
 * rows: optional row-id list (selection) or 0 for 0..n-1.
 * Returns group count; out_rep[g] = representative row, out_res[g*E+e]. */
orc_u64 orc_aggregate_cpu_reduce(orc_u64 n, const orc_u64* rows, const orc_f64* const* keys,
                                 orc_u32 key_count, const orc_u32* funcs,
                                 const orc_f64* const* args, orc_u32 expr_count,
                                 orc_u64* out_rep, orc_f64* out_res);

/* Typed wrapper: GROUP BY one column over typed columns, optional WHERE
 * (IR) -- loads values as f64 the way qe_row_load_value does and calls
 * orc_aggregate_cpu_reduce.  arg_cols[e] < 0 means COUNT(*). */
orc_u64 orc_group_aggregate(const orc_column* cols, orc_u64 rows, const orc_ir_node* where,
                            orc_u32 where_count, const int* key_cols, orc_u32 key_count, const orc_u32* funcs,
                            const int* arg_cols, orc_u32 expr_count, orc_f64* out_keys,
                            orc_f64* out_res, orc_u64 max_groups);
/* qe_aggregate_check_sum_exact (query_exec.c:7028-7049): 0 when a SUM over
 * a whole-number argument type reached 2^53 */
int orc_check_sum_exact(const orc_u32* funcs, const orc_u32* arg_types, orc_u32 expr_count, const orc_f64* res,
                        orc_u64 groups);

/* Join + aggregate, the semantics of qe_aggregate_join_pushdown
 * (query_exec.c:8442-8775) and of the general path it stands in for
 * (qe_hash_join_impl :14541 then qe_aggregate): every fact row whose key
 * equals the key of a dim row that passes the dim filter contributes to the
 * group of that dim row's group columns (group_count of them, 0..4: a
 * group is one combination of their values; 0 columns: one group); group
 * keys come out group_count values per group.
 * Fact-side aggregates over fact_arg columns (0 = COUNT).  Groups in first-
 * appearance order over fact rows.  Returns group count. */
orc_u64 orc_join_aggregate(orc_u64 fact_rows, const orc_u32* fact_key, const orc_u32* const* fact_args,
                           const orc_u32* funcs, orc_u32 expr_count, orc_u64 dim_rows,
                           const orc_u32* dim_key, const orc_u8* dim_pass, const orc_u32* const* dim_groups,
                           orc_u32 group_count, orc_f64* out_group_keys, orc_f64* out_res, orc_u64 max_groups,
                           int threads);

/* multithreaded numeric-AND bitmask count (qe_scan_filter_chunked_cpu,
 * query_exec.c:1240-1367, count_only) for leaves col op target (op 0..4).
 * Used as the timed CPU baseline of a filter count. */
orc_u64 orc_hash_join(const orc_column* probe_key, const orc_u64* probe_rows, orc_u64 np,
                      const orc_column* build_key, const orc_u64* build_rows, orc_u64 nb, orc_u32 join_type,
                      orc_u64* out, orc_u64 cap);
orc_u64 orc_scan_chunked_count(const orc_column* cols, const orc_u32* leaf_col, const orc_u32* leaf_op,
                               const orc_f64* leaf_target, orc_u32 leaf_count, orc_u64 rows, int workers);

/* dense GROUP BY on the CPU, multithreaded, exact integer accumulators;
 * used as the CPU baseline for grouped aggregation at scale (results equal
 * orc_aggregate_cpu_reduce after key-order normalisation for integer data).
 * out arrays are indexed by key - key_min (range entries). */
void orc_dense_group_u32(const orc_u32* key, const orc_u32* arg, orc_u64 rows, orc_u32 key_min,
                         orc_u32 range, int workers, orc_u64* out_count, orc_u64* out_sum,
                         orc_u32* out_min, orc_u32* out_max);

int orc_hw_threads(void);

#ifdef __cplusplus
}
#endif
#endif
