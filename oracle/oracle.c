/*
 * oracle.c - CPU restatement of ComputeQL's CPU operator path (scan/filter,
 * global and grouped aggregates, join + aggregate) plus the seeded and
 * closed-form table generators.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker / timed CPU baseline.  The
 * product library never links or calls it.
 *
 * Reference files (relative to the ComputeQL tree):
 *   src/query_exec/query_exec.c   operator semantics (cited per function)
 *   src/tests/bench_common.h      splitmix64 (:30-37), seed (:59)
 *   src/tests/scale/bench_select_scale.c  closed-form table and answers
 *   src/thread_pool/thread_pool.c tp_divide_work (:236-249)
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define ORC_GOLDEN 0x9E3779B97F4A7C15ULL
#define ORC_FK_SALT 0x5851F42D4C957F2DULL
#define ORC_PERM_SALT 0x2545F4914F6CDD1DULL
#define ORC_ATTR_SALT 0x14057B7EF767814FULL

int orc_hw_threads(void)
{
  long n = sysconf(_SC_NPROCESSORS_ONLN);
  return n > 0 ? (int)n : 1;
}

/* ------------------------------------------------------------------ */
/* parallel-for over task ids (stands in for tp_for_parallel)          */
/* ------------------------------------------------------------------ */
typedef void (*orc_task_fn)(void* ctx, orc_u64 task_id);
typedef struct {
  orc_task_fn fn;
  void* ctx;
  orc_u64 task_count;
  _Atomic orc_u64 next;
} orc_pool_job;

static void* orc_pool_worker(void* raw)
{
  orc_pool_job* job = (orc_pool_job*)raw;
  for (;;) {
    orc_u64 t = atomic_fetch_add(&job->next, 1);
    if (t >= job->task_count) break;
    job->fn(job->ctx, t);
  }
  return 0;
}

static void orc_parallel_for(orc_u64 task_count, int threads, orc_task_fn fn, void* ctx)
{
  if (threads < 1) threads = 1;
  if ((orc_u64)threads > task_count) threads = (int)task_count;
  orc_pool_job job;
  job.fn = fn;
  job.ctx = ctx;
  job.task_count = task_count;
  atomic_store(&job.next, 0);
  if (threads <= 1) {
    orc_pool_worker(&job);
    return;
  }
  pthread_t* tids = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  int started = 0;
  for (int i = 0; i < threads - 1; i++) {
    if (pthread_create(&tids[i], 0, orc_pool_worker, &job) == 0) started++;
    else break;
  }
  orc_pool_worker(&job);
  for (int i = 0; i < started; i++) pthread_join(tids[i], 0);
  free(tids);
}

/* tp_divide_work, thread_pool.c:236-249: ceil split, contiguous ranges */
static void orc_divide_work(orc_u64 item_count, orc_u64 workers, orc_u64 t, orc_u64* lo, orc_u64* hi)
{
  orc_u64 per = (item_count + workers - 1) / workers;
  orc_u64 a = t * per, b = t * per + per;
  *lo = a < item_count ? a : item_count;
  *hi = b < item_count ? b : item_count;
}

/* ------------------------------------------------------------------ */
/* generators                                                           */
/* ------------------------------------------------------------------ */
/* bench_rng_next, bench_common.h:30-37, addressed by output index */
orc_u64 orc_splitmix64_at(orc_u64 seed, orc_u64 k)
{
  orc_u64 z = seed + (k + 1) * ORC_GOLDEN;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

typedef struct {
  orc_u64 seed, row0, n, tasks;
  orc_u32 *id, *region, *amount, *ts;
  const orc_u32* dim_keys;
  orc_u64 dim_rows;
  orc_u32* fk;
} orc_gen_ctx;

static void orc_gen_events_task(void* raw, orc_u64 t)
{
  orc_gen_ctx* c = (orc_gen_ctx*)raw;
  orc_u64 lo, hi;
  orc_divide_work(c->n, c->tasks, t, &lo, &hi);
  for (orc_u64 i = lo; i < hi; i++) {
    orc_u64 r = c->row0 + i;
    if (c->id) c->id[i] = (orc_u32)r;
    if (c->region) c->region[i] = (orc_u32)(orc_splitmix64_at(c->seed, 3 * r + 0) % 50u);
    if (c->amount) c->amount[i] = (orc_u32)(orc_splitmix64_at(c->seed, 3 * r + 1) % 10000u);
    if (c->ts) c->ts[i] = (orc_u32)(orc_splitmix64_at(c->seed, 3 * r + 2) >> 32);
  }
}

void orc_gen_events(orc_u64 seed, orc_u64 row0, orc_u64 n, orc_u32* id, orc_u32* region,
                    orc_u32* amount, orc_u32* ts, int threads)
{
  orc_gen_ctx c;
  memset(&c, 0, sizeof(c));
  c.seed = seed; c.row0 = row0; c.n = n;
  c.id = id; c.region = region; c.amount = amount; c.ts = ts;
  c.tasks = n < 65536 ? 1 : (orc_u64)(threads < 1 ? 1 : threads) * 4;
  orc_parallel_for(c.tasks, threads, orc_gen_events_task, &c);
}

static void orc_gen_fk_task(void* raw, orc_u64 t)
{
  orc_gen_ctx* c = (orc_gen_ctx*)raw;
  orc_u64 lo, hi;
  orc_divide_work(c->n, c->tasks, t, &lo, &hi);
  orc_u64 s = c->seed ^ ORC_FK_SALT;
  for (orc_u64 i = lo; i < hi; i++) {
    orc_u64 r = c->row0 + i;
    c->fk[i] = c->dim_keys[orc_splitmix64_at(s, r) % c->dim_rows];
  }
}

void orc_gen_fact_fk(orc_u64 seed, orc_u64 row0, orc_u64 n, const orc_u32* dim_keys,
                     orc_u64 dim_rows, orc_u32* fk, int threads)
{
  orc_gen_ctx c;
  memset(&c, 0, sizeof(c));
  c.seed = seed; c.row0 = row0; c.n = n; c.dim_keys = dim_keys; c.dim_rows = dim_rows; c.fk = fk;
  c.tasks = n < 65536 ? 1 : (orc_u64)(threads < 1 ? 1 : threads) * 4;
  orc_parallel_for(c.tasks, threads, orc_gen_fk_task, &c);
}

void orc_gen_dim(orc_u64 seed, orc_u64 n, orc_u32 attr_range, int spread_keys,
                 orc_u32* keys, orc_u32* attr)
{
  for (orc_u64 i = 0; i < n; i++) keys[i] = (orc_u32)i;
  orc_u64 ps = seed ^ ORC_PERM_SALT;
  for (orc_u64 i = n; i > 1; i--) {
    orc_u64 j = orc_splitmix64_at(ps, i - 1) % i;
    orc_u32 tmp = keys[i - 1];
    keys[i - 1] = keys[j];
    keys[j] = tmp;
  }
  if (spread_keys) {
    for (orc_u64 i = 0; i < n; i++) keys[i] = keys[i] * 0x85EBCA77u + 0x165667B1u;
  }
  orc_u64 as = seed ^ ORC_ATTR_SALT;
  if (attr) {
    for (orc_u64 i = 0; i < n; i++) attr[i] = (orc_u32)(orc_splitmix64_at(as, i) % attr_range);
  }
}

/* bench_select_scale.c:39-45 and ss_fill_columns :172-197 */
void orc_gen_scale_events(orc_u64 row0, orc_u64 n, orc_u32* id, orc_u32* region,
                          orc_u32* amount, orc_u32* needle_key)
{
  for (orc_u64 i = 0; i < n; i++) {
    orc_u64 r = row0 + i;
    if (id) id[i] = (orc_u32)(r + 1);
    if (region) region[i] = (orc_u32)(r % 50);
    if (amount) amount[i] = (orc_u32)(r % 10000);
    if (needle_key) needle_key[i] = orc_ss_key_of_row(r);
  }
}

/* bench_select_scale.c:98-105 */
orc_f64 orc_ss_sum_amount(orc_u64 rows)
{
  orc_u64 full = rows / 10000, rem = rows % 10000;
  orc_u64 cycle = (orc_u64)10000 * 9999 / 2;
  return (orc_f64)(full * cycle + (rem > 0 ? rem * (rem - 1) / 2 : 0));
}

/* bench_select_scale.c:107-115 */
orc_u64 orc_ss_count_amount_above(orc_u64 rows, orc_u64 threshold)
{
  orc_u64 full = rows / 10000, rem = rows % 10000;
  orc_u64 per = 10000 - 1 - threshold;
  orc_u64 partial = (rem > threshold + 1) ? (rem - threshold - 1) : 0;
  return full * per + partial;
}

/* bench_select_scale.c:117-121 / :124-128 */
orc_u64 orc_ss_count_region(orc_u64 rows, orc_u64 region)
{
  return rows / 50 + ((rows % 50) > region ? 1 : 0);
}

/* bench_select_scale.c:130-139 */
orc_f64 orc_ss_region_amount_sum(orc_u64 rows, orc_u64 region)
{
  orc_u64 n = orc_ss_count_region(rows, region);
  orc_u64 cycles = n / 200, rem = n % 200;
  orc_u64 cycle_sum = 200 * 199 / 2;
  orc_u64 partial = rem > 0 ? rem * (rem - 1) / 2 : 0;
  return (orc_f64)(n * region + 50 * (cycles * cycle_sum + partial));
}

/* bench_select_scale.c:141-145 */
orc_u64 orc_ss_amount_rows(orc_u64 rows, orc_u64 amount)
{
  return rows / 10000 + ((rows % 10000) > amount ? 1 : 0);
}

/* bench_select_scale.c:147-154 */
orc_f64 orc_ss_weight_sum(orc_u64 count)
{
  orc_u64 cycles = count / 100, rem = count % 100;
  orc_u64 partial = rem > 0 ? rem * (rem - 1) / 2 : 0;
  return (orc_f64)(cycles * (100 * 99 / 2) + partial);
}

/* bench_select_scale.c:156-160 */
orc_u32 orc_ss_key_of_row(orc_u64 row)
{
  return (orc_u32)(row * 2654435761ULL);
}

/* ------------------------------------------------------------------ */
/* row predicate semantics                                              */
/* ------------------------------------------------------------------ */
static orc_f64 orc_load(const orc_column* col, orc_u64 row)
{
  switch (col->type) {
    case ORC_TYPE_U32: return (orc_f64)((const orc_u32*)col->data)[row];
    case ORC_TYPE_I32: return (orc_f64)((const orc_s32*)col->data)[row];
    case ORC_TYPE_U64: return (orc_f64)((const orc_u64*)col->data)[row];
    case ORC_TYPE_I64: return (orc_f64)((const orc_s64*)col->data)[row];
    case ORC_TYPE_F32: return (orc_f64)((const float*)col->data)[row];
    case ORC_TYPE_F64: return ((const orc_f64*)col->data)[row];
    default: return 0.0;
  }
}

static orc_u32 orc_subtree_end(const orc_ir_node* nodes, orc_u32 i)
{
  orc_u32 end = i + 1;
  for (orc_u32 c = 0; c < nodes[i].child_count; c++) end = orc_subtree_end(nodes, end);
  return end;
}

/* qe_row_load_value (numeric): column value as f64 or the literal */
static orc_f64 orc_value(const orc_ir_node* n, const orc_column* cols, orc_u64 row, int* is_null)
{
  *is_null = 0;
  if (n->type == ORC_IR_COLUMN) {
    const orc_column* c = &cols[n->column];
    if (c->null_flags && c->null_flags[row]) {
      *is_null = 1;
      return 0.0;
    }
    return orc_load(c, row);
  }
  return n->value;
}

static int orc_eval_at(const orc_ir_node* nodes, orc_u32 i, const orc_column* cols, orc_u64 row)
{
  const orc_ir_node* n = &nodes[i];
  if (n->type != ORC_IR_OPERATOR) {
    /* query_exec.c:13878-13887: bare value, nonzero = true, NULL = false */
    int is_null = 0;
    orc_f64 v = orc_value(n, cols, row, &is_null);
    return is_null ? 0 : (v != 0.0);
  }
  orc_u32 left = i + 1;
  orc_u32 right = n->child_count > 1 ? orc_subtree_end(nodes, left) : 0;
  switch (n->op) {
    case ORC_OP_AND: /* :13893-13896 */
      return orc_eval_at(nodes, left, cols, row) && orc_eval_at(nodes, right, cols, row);
    case ORC_OP_OR: /* :13897-13900 */
      return orc_eval_at(nodes, left, cols, row) || orc_eval_at(nodes, right, cols, row);
    case ORC_OP_IS_NULL:
    case ORC_OP_IS_NOT_NULL: { /* :13902-13917 */
      int lnull = 0;
      orc_value(&nodes[left], cols, row, &lnull);
      return n->op == ORC_OP_IS_NOT_NULL ? !lnull : lnull;
    }
    case ORC_OP_IN:
    case ORC_OP_NOT_IN: { /* :13919-13935: NULL [NOT] IN (...) is never true */
      int lnull = 0;
      orc_f64 lv = orc_value(&nodes[left], cols, row, &lnull);
      if (lnull) return 0;
      const orc_ir_node* list = &nodes[right];
      int found = 0;
      for (orc_u32 k = 0; k < list->child_count; k++) {
        if (nodes[right + 1 + k].value == lv) {
          found = 1;
          break;
        }
      }
      return n->op == ORC_OP_NOT_IN ? !found : found;
    }
    default: break;
  }
  /* :13971-14043: numeric comparison, a NULL operand is never true */
  int lnull = 0, rnull = 0;
  orc_f64 lv = orc_value(&nodes[left], cols, row, &lnull);
  orc_f64 rv = orc_value(&nodes[right], cols, row, &rnull);
  if (lnull || rnull) return 0;
  switch (n->op) {
    case ORC_OP_EQ: return lv == rv;
    case ORC_OP_NE: return lv != rv;
    case ORC_OP_LE: return lv <= rv;
    case ORC_OP_GE: return lv >= rv;
    case ORC_OP_LT: return lv < rv;
    case ORC_OP_GT: return lv > rv;
    default: return 1;
  }
}

int orc_row_condition_eval(const orc_ir_node* nodes, orc_u32 node_count, const orc_column* cols,
                           orc_u64 row)
{
  if (!nodes || node_count == 0) return 1;
  return orc_eval_at(nodes, 0, cols, row);
}

typedef struct {
  const orc_ir_node* nodes;
  orc_u32 node_count;
  const orc_column* cols;
  orc_u64 rows, tasks;
  orc_u64* counts;
  orc_u64* offsets;
  orc_u64* out;
  orc_u64 cap;
} orc_scan_ctx;

static void orc_scan_count_task(void* raw, orc_u64 t)
{
  orc_scan_ctx* c = (orc_scan_ctx*)raw;
  orc_u64 lo, hi, k = 0;
  orc_divide_work(c->rows, c->tasks, t, &lo, &hi);
  for (orc_u64 i = lo; i < hi; i++) k += (orc_u64)orc_row_condition_eval(c->nodes, c->node_count, c->cols, i);
  c->counts[t] = k;
}

static void orc_scan_write_task(void* raw, orc_u64 t)
{
  orc_scan_ctx* c = (orc_scan_ctx*)raw;
  orc_u64 lo, hi, k = c->offsets[t];
  orc_divide_work(c->rows, c->tasks, t, &lo, &hi);
  for (orc_u64 i = lo; i < hi; i++) {
    if (orc_row_condition_eval(c->nodes, c->node_count, c->cols, i)) {
      if (k < c->cap) c->out[k] = i;
      k++;
    }
  }
}

/* qe_cpu_scan_filter, query_exec.c:14064-14135: tasks over contiguous
 * ranges, matches concatenated in task order = row order */
orc_u64 orc_cpu_scan_filter(const orc_ir_node* nodes, orc_u32 node_count, const orc_column* cols,
                            orc_u64 rows, orc_u64* out, orc_u64 cap, int threads)
{
  if (rows == 0) return 0;
  orc_scan_ctx c;
  memset(&c, 0, sizeof(c));
  c.nodes = nodes; c.node_count = node_count; c.cols = cols; c.rows = rows;
  c.tasks = rows > 1 ? (orc_u64)(threads < 1 ? 1 : threads) : 1;
  if (c.tasks > rows) c.tasks = rows;
  c.counts = (orc_u64*)calloc(c.tasks, sizeof(orc_u64));
  c.offsets = (orc_u64*)calloc(c.tasks, sizeof(orc_u64));
  orc_parallel_for(c.tasks, threads, orc_scan_count_task, &c);
  orc_u64 total = 0;
  for (orc_u64 t = 0; t < c.tasks; t++) {
    c.offsets[t] = total;
    total += c.counts[t];
  }
  if (out) {
    c.out = out;
    c.cap = cap;
    orc_parallel_for(c.tasks, threads, orc_scan_write_task, &c);
  }
  free(c.counts);
  free(c.offsets);
  return total;
}

/* qe_scan_int_interval, query_exec.c:1378-1404 */
int orc_scan_int_interval(orc_u32 op, orc_f64 target, orc_f64 code_offset, orc_f64 max_code,
                          orc_u32* out_lo, orc_u32* out_span)
{
  if (target != target) return 0;
  orc_f64 low = -1e18, high = 1e18;
  switch (op) {
    case 0: if (floor(target) != target) return 0; low = target; high = target; break;
    case 1: high = ceil(target) - 1.0; break;
    case 2: high = floor(target); break;
    case 3: low = floor(target) + 1.0; break;
    case 4: low = ceil(target); break;
    default: return 0;
  }
  low = fmax(low - code_offset, 0.0);
  high = fmin(high - code_offset, max_code);
  if (low > high) return 0;
  *out_lo = (orc_u32)low;
  *out_span = (orc_u32)(high - low);
  return 1;
}

/* ------------------------------------------------------------------ */
/* global aggregate                                                     */
/* ------------------------------------------------------------------ */
typedef struct {
  const orc_column* col;
  orc_u64 base, rows, tasks;
  orc_f64 *sum, *mn, *mx;
} orc_reduce_ctx;

/* qe_chunk_reduce_task + QE_CHUNK_REDUCE_LOOP, query_exec.c:7472-7500 */
static void orc_chunk_reduce_task(void* raw, orc_u64 t)
{
  orc_reduce_ctx* c = (orc_reduce_ctx*)raw;
  orc_u64 lo, hi;
  orc_divide_work(c->rows, c->tasks, t, &lo, &hi);
  orc_f64 sum = 0.0, mn = DBL_MAX, mx = -DBL_MAX;
#define ORC_LOOP(T)                                         \
  {                                                         \
    const T* v = (const T*)c->col->data + c->base;          \
    for (orc_u64 i = lo; i < hi; i++) {                     \
      orc_f64 x = (orc_f64)v[i];                            \
      sum += x;                                             \
      if (x < mn) mn = x;                                   \
      if (x > mx) mx = x;                                   \
    }                                                       \
  }
  switch (c->col->type) {
    case ORC_TYPE_U32: ORC_LOOP(orc_u32) break;
    case ORC_TYPE_I32: ORC_LOOP(orc_s32) break;
    case ORC_TYPE_U64: ORC_LOOP(orc_u64) break;
    case ORC_TYPE_I64: ORC_LOOP(orc_s64) break;
    case ORC_TYPE_F32: ORC_LOOP(float) break;
    case ORC_TYPE_F64: ORC_LOOP(orc_f64) break;
    default: break;
  }
#undef ORC_LOOP
  c->sum[t] = sum;
  c->mn[t] = mn;
  c->mx[t] = mx;
}

void orc_aggregate_global_chunked(const orc_column* cols, orc_u32 col_count, orc_u64 rows,
                                  int workers, orc_f64* out_sum, orc_f64* out_min, orc_f64* out_max)
{
  const orc_u64 chunk_rows = 67108864; /* QE_AGG_CHUNK_ROWS default, :9159 */
  if (workers < 1) workers = 1;
  for (orc_u32 c = 0; c < col_count; c++) {
    out_sum[c] = 0.0;
    out_min[c] = DBL_MAX;
    out_max[c] = -DBL_MAX;
  }
  orc_u64 max_tasks = (orc_u64)workers;
  orc_f64* ts = (orc_f64*)malloc(sizeof(orc_f64) * max_tasks);
  orc_f64* tn = (orc_f64*)malloc(sizeof(orc_f64) * max_tasks);
  orc_f64* tx = (orc_f64*)malloc(sizeof(orc_f64) * max_tasks);
  for (orc_u64 start = 0; start < rows; start += chunk_rows) {
    orc_u64 end = start + chunk_rows < rows ? start + chunk_rows : rows;
    orc_u64 n = end - start;
    orc_u64 tasks = n / 65536;
    if (tasks > (orc_u64)workers) tasks = (orc_u64)workers;
    if (tasks < 1) tasks = 1;
    for (orc_u32 c = 0; c < col_count; c++) {
      orc_reduce_ctx ctx = {&cols[c], start, n, tasks, ts, tn, tx};
      orc_parallel_for(tasks, workers, orc_chunk_reduce_task, &ctx);
      for (orc_u64 t = 0; t < tasks; t++) { /* :9196-9207, task order */
        out_sum[c] += ts[t];
        if (tn[t] < out_min[c]) out_min[c] = tn[t];
        if (tx[t] > out_max[c]) out_max[c] = tx[t];
      }
    }
  }
  free(ts);
  free(tn);
  free(tx);
}

/* ------------------------------------------------------------------ */
/* grouped aggregate                                                    */
/* ------------------------------------------------------------------ */
/* qe_aggregate_hash_key, query_exec.c:7201-7234 (numeric keys) */
static orc_u64 orc_hash_keys(const orc_f64* const* keys, orc_u32 key_count, orc_u64 row)
{
  orc_u64 hash = 14695981039346656037ULL;
  for (orc_u32 c = 0; c < key_count; c++) {
    orc_f64 v = keys[c][row];
    if (v == 0.0) v = 0.0;
    orc_u64 bits;
    memcpy(&bits, &v, sizeof(bits));
    hash ^= bits;
    hash *= 0x9E3779B97F4A7C15ULL;
    hash ^= hash >> 32;
  }
  return hash;
}

static int orc_keys_equal(const orc_f64* const* keys, orc_u32 key_count, orc_u64 a, orc_u64 b)
{
  for (orc_u32 c = 0; c < key_count; c++)
    if (keys[c][a] != keys[c][b]) return 0;
  return 1;
}

typedef struct { orc_f64 sum; orc_u64 count; orc_f64 mn, mx; } orc_acc;

/* qe_aggregate_finish, query_exec.c:7269-7293 */
static orc_f64 orc_finish(orc_u32 func, const orc_acc* a)
{
  switch (func) {
    case ORC_AGG_COUNT: return (orc_f64)a->count;
    case ORC_AGG_SUM: return a->sum;
    case ORC_AGG_AVG: return a->count > 0 ? a->sum / (orc_f64)a->count : 0.0;
    case ORC_AGG_MIN: return a->mn;
    default: return a->mx;
  }
}

/* qe_aggregate_cpu_reduce, query_exec.c:7297-7399.  Keys/args are indexed
 * by position 0..n-1 in the (already gathered) input. */
orc_u64 orc_aggregate_cpu_reduce(orc_u64 n, const orc_u64* rows, const orc_f64* const* keys,
                                 orc_u32 key_count, const orc_u32* funcs,
                                 const orc_f64* const* args, orc_u32 expr_count,
                                 orc_u64* out_rep, orc_f64* out_res)
{
  (void)rows;
  orc_u64 stride = expr_count > 0 ? expr_count : 1;
  orc_u64 max_groups = key_count == 0 ? 1 : (n > 0 ? n : 1);
  orc_u64 table_size = 16;
  while (table_size < max_groups * 2) table_size <<= 1;
  orc_u32* table = (orc_u32*)malloc(sizeof(orc_u32) * table_size);
  for (orc_u64 s = 0; s < table_size; s++) table[s] = 0xffffffffu;
  orc_u64* reps = (orc_u64*)malloc(sizeof(orc_u64) * max_groups);
  orc_acc* acc = (orc_acc*)malloc(sizeof(orc_acc) * max_groups * stride);
  orc_u64 groups = 0;
  for (orc_u64 row = 0; row < n; row++) {
    orc_u64 g = 0;
    int fresh = 0;
    if (key_count == 0) {
      fresh = row == 0;
    } else {
      orc_u64 slot = orc_hash_keys(keys, key_count, row) & (table_size - 1);
      for (;;) {
        orc_u32 occ = table[slot];
        if (occ == 0xffffffffu) {
          g = groups;
          table[slot] = (orc_u32)g;
          fresh = 1;
          break;
        }
        if (orc_keys_equal(keys, key_count, reps[occ], row)) {
          g = occ;
          break;
        }
        slot = (slot + 1) & (table_size - 1);
      }
    }
    if (fresh) {
      reps[g] = row;
      for (orc_u64 e = 0; e < stride; e++) {
        orc_acc* a = &acc[g * stride + e];
        a->sum = 0.0;
        a->count = 0;
        a->mn = 1.0e300;
        a->mx = -1.0e300;
      }
      groups++;
    }
    for (orc_u32 e = 0; e < expr_count; e++) {
      orc_acc* a = &acc[g * stride + e];
      orc_f64 v = args[e] ? args[e][row] : 0.0;
      a->sum += v;
      a->count += 1;
      if (v < a->mn) a->mn = v;
      if (v > a->mx) a->mx = v;
    }
  }
  for (orc_u64 g = 0; g < groups; g++) {
    if (out_rep) out_rep[g] = reps[g];
    for (orc_u32 e = 0; e < expr_count; e++) out_res[g * expr_count + e] = orc_finish(funcs[e], &acc[g * stride + e]);
  }
  free(table);
  free(reps);
  free(acc);
  return groups;
}

/* qe_aggregate_impl's GROUP BY over 0..4 key columns (query_exec.c:
 * 10280-10299 gathers each key to f64 and hashes the tuple, :7201-7234) */
orc_u64 orc_group_aggregate(const orc_column* cols, orc_u64 rows, const orc_ir_node* where,
                            orc_u32 where_count, const int* key_cols, orc_u32 key_count, const orc_u32* funcs,
                            const int* arg_cols, orc_u32 expr_count, orc_f64* out_keys,
                            orc_f64* out_res, orc_u64 max_groups)
{
  orc_u64* sel = 0;
  orc_u64 n = rows;
  if (where && where_count) {
    sel = (orc_u64*)malloc(sizeof(orc_u64) * (rows ? rows : 1));
    n = orc_cpu_scan_filter(where, where_count, cols, rows, sel, rows, 1);
  }
  orc_f64* keyv[4] = {0, 0, 0, 0};
  if (key_count > 4) key_count = 4;
  for (orc_u32 k = 0; k < key_count; k++) {
    keyv[k] = (orc_f64*)malloc(sizeof(orc_f64) * (n ? n : 1));
    for (orc_u64 i = 0; i < n; i++) keyv[k][i] = orc_load(&cols[key_cols[k]], sel ? sel[i] : i);
  }
  orc_f64** argv = (orc_f64**)calloc(expr_count ? expr_count : 1, sizeof(orc_f64*));
  for (orc_u32 e = 0; e < expr_count; e++) {
    if (arg_cols[e] < 0) continue;
    argv[e] = (orc_f64*)malloc(sizeof(orc_f64) * (n ? n : 1));
    for (orc_u64 i = 0; i < n; i++) argv[e][i] = orc_load(&cols[arg_cols[e]], sel ? sel[i] : i);
  }
  orc_u64* reps = (orc_u64*)malloc(sizeof(orc_u64) * (n ? n : 1));
  orc_f64* res = (orc_f64*)malloc(sizeof(orc_f64) * (n ? n : 1) * (expr_count ? expr_count : 1));
  const orc_f64* kp[4] = {keyv[0], keyv[1], keyv[2], keyv[3]};
  orc_u64 g = orc_aggregate_cpu_reduce(n, 0, kp, key_count, funcs, (const orc_f64* const*)argv, expr_count, reps, res);
  /* an ungrouped aggregate over no rows still yields one row of zeros
   * (query_exec.c:10335-10346) */
  if (key_count == 0 && g == 0) {
    g = 1;
    for (orc_u32 e = 0; e < expr_count; e++) res[e] = 0.0;
    reps[0] = 0;
  }
  orc_u64 out_g = g < max_groups ? g : max_groups;
  for (orc_u64 i = 0; i < out_g; i++) {
    for (orc_u32 k = 0; out_keys && k < key_count; k++) out_keys[i * key_count + k] = n ? keyv[k][reps[i]] : 0.0;
    for (orc_u32 e = 0; e < expr_count; e++) out_res[i * expr_count + e] = res[i * expr_count + e];
  }
  free(sel);
  for (orc_u32 k = 0; k < 4; k++) free(keyv[k]);
  for (orc_u32 e = 0; e < expr_count; e++) free(argv[e]);
  free(argv);
  free(reps);
  free(res);
  return g;
}

/* ------------------------------------------------------------------ */
/* join + aggregate                                                     */
/* ------------------------------------------------------------------ */
typedef struct {
  orc_u64 fact_rows, tasks;
  const orc_u32* fact_key;
  const orc_u32* const* fact_args;
  const orc_u32* funcs;
  orc_u32 expr_count;
  const orc_u32* dim_key;
  const orc_u8* dim_pass;
  orc_u32* heads; /* dim hash: slot -> first dim row + 1 */
  orc_u32* next;  /* dim row -> next dim row + 1 with the same slot */
  orc_u64 hash_mask;
  /* group ids: dense over the distinct dim group values */
  orc_u32* group_of_dim; /* dim row -> group index (0 if no grouping) */
  orc_u64 group_count;
  /* per task: acc[task][group][expr], first fact row per group */
  orc_acc* acc;
  orc_u64* first_row;
} orc_join_ctx;

static orc_u64 orc_dim_slot(orc_u32 key, orc_u64 mask)
{
  orc_u64 h = (orc_u64)key * 0x9E3779B97F4A7C15ULL;
  return (h >> 29) & mask;
}

static void orc_join_task(void* raw, orc_u64 t)
{
  orc_join_ctx* c = (orc_join_ctx*)raw;
  orc_u64 lo, hi;
  orc_divide_work(c->fact_rows, c->tasks, t, &lo, &hi);
  orc_u64 E = c->expr_count;
  orc_acc* acc = c->acc + t * c->group_count * E;
  orc_u64* first = c->first_row + t * c->group_count;
  for (orc_u64 r = lo; r < hi; r++) {
    orc_u32 key = c->fact_key[r];
    for (orc_u32 d1 = c->heads[orc_dim_slot(key, c->hash_mask)]; d1; d1 = c->next[d1 - 1]) {
      orc_u32 d = d1 - 1;
      if (c->dim_key[d] != key) continue;
      if (c->dim_pass && !c->dim_pass[d]) continue;
      orc_u32 g = c->group_of_dim[d];
      if (first[g] == ~(orc_u64)0) first[g] = r;
      for (orc_u64 e = 0; e < E; e++) {
        orc_acc* a = &acc[g * E + e];
        orc_f64 v = c->fact_args[e] ? (orc_f64)c->fact_args[e][r] : 0.0;
        a->sum += v;
        a->count += 1;
        if (v < a->mn) a->mn = v;
        if (v > a->mx) a->mx = v;
      }
    }
  }
}

/* the dim group columns of dim row d as one tuple (unused columns 0) */
static void orc_dim_tuple(const orc_u32* const* dim_groups, orc_u32 group_count, orc_u64 d, orc_u32* out)
{
  for (orc_u32 k = 0; k < 4; k++) out[k] = k < group_count ? dim_groups[k][d] : 0u;
}

static orc_u64 orc_tuple_slot(const orc_u32* v, orc_u64 mask)
{
  orc_u64 h = 0;
  for (orc_u32 k = 0; k < 4; k++) h = (h ^ v[k]) * 0x9E3779B97F4A7C15ULL;
  return (h >> 29) & mask;
}

orc_u64 orc_join_aggregate(orc_u64 fact_rows, const orc_u32* fact_key, const orc_u32* const* fact_args,
                           const orc_u32* funcs, orc_u32 expr_count, orc_u64 dim_rows,
                           const orc_u32* dim_key, const orc_u8* dim_pass, const orc_u32* const* dim_groups,
                           orc_u32 group_count, orc_f64* out_group_keys, orc_f64* out_res, orc_u64 max_groups,
                           int threads)
{
  orc_join_ctx c;
  memset(&c, 0, sizeof(c));
  if (group_count > 4) group_count = 4;
  c.fact_rows = fact_rows; c.fact_key = fact_key; c.fact_args = fact_args; c.funcs = funcs;
  c.expr_count = expr_count; c.dim_key = dim_key; c.dim_pass = dim_pass;
  orc_u64 size = 64;
  while (size < dim_rows * 2 + 2) size <<= 1;
  c.hash_mask = size - 1;
  c.heads = (orc_u32*)calloc(size, sizeof(orc_u32));
  c.next = (orc_u32*)calloc(dim_rows ? dim_rows : 1, sizeof(orc_u32));
  /* insert in reverse so chains list dim rows in ascending order */
  for (orc_u64 d = dim_rows; d > 0; d--) {
    orc_u64 s = orc_dim_slot(dim_key[d - 1], c.hash_mask);
    c.next[d - 1] = c.heads[s];
    c.heads[s] = (orc_u32)d;
  }
  /* distinct group tuples of the dim rows (any order; output order is
   * fixed below by first fact appearance): a group is a distinct
   * combination of the group column values (query_exec.c:8621-8690) */
  c.group_of_dim = (orc_u32*)calloc(dim_rows ? dim_rows : 1, sizeof(orc_u32));
  orc_u32* group_value = (orc_u32*)calloc(4 * (dim_rows ? dim_rows : 1), sizeof(orc_u32));
  c.group_count = 1;
  if (group_count) {
    orc_u64 gsize = 64;
    while (gsize < dim_rows * 2 + 2) gsize <<= 1;
    orc_u32* gslots = (orc_u32*)calloc(gsize, sizeof(orc_u32));
    c.group_count = 0;
    for (orc_u64 d = 0; d < dim_rows; d++) {
      orc_u32 v[4];
      orc_dim_tuple(dim_groups, group_count, d, v);
      orc_u64 s = orc_tuple_slot(v, gsize - 1);
      for (;;) {
        if (gslots[s] == 0) {
          memcpy(&group_value[4 * c.group_count], v, sizeof(v));
          gslots[s] = (orc_u32)(++c.group_count);
          c.group_of_dim[d] = (orc_u32)(c.group_count - 1);
          break;
        }
        if (!memcmp(&group_value[4 * (gslots[s] - 1)], v, sizeof(v))) {
          c.group_of_dim[d] = gslots[s] - 1;
          break;
        }
        s = (s + 1) & (gsize - 1);
      }
    }
    free(gslots);
  }
  if (threads < 1) threads = 1;
  c.tasks = fact_rows < 65536 ? 1 : (orc_u64)threads * 4;
  orc_u64 E = expr_count;
  c.acc = (orc_acc*)malloc(sizeof(orc_acc) * c.tasks * c.group_count * (E ? E : 1));
  c.first_row = (orc_u64*)malloc(sizeof(orc_u64) * c.tasks * c.group_count);
  for (orc_u64 i = 0; i < c.tasks * c.group_count * E; i++) {
    c.acc[i].sum = 0.0;
    c.acc[i].count = 0;
    c.acc[i].mn = 1.0e300;
    c.acc[i].mx = -1.0e300;
  }
  for (orc_u64 i = 0; i < c.tasks * c.group_count; i++) c.first_row[i] = ~(orc_u64)0;
  orc_parallel_for(c.tasks, threads, orc_join_task, &c);
  /* fold tasks in task order; order groups by first fact row */
  orc_u64* first = (orc_u64*)malloc(sizeof(orc_u64) * c.group_count);
  orc_acc* tot = (orc_acc*)malloc(sizeof(orc_acc) * c.group_count * (E ? E : 1));
  for (orc_u64 g = 0; g < c.group_count; g++) {
    first[g] = ~(orc_u64)0;
    for (orc_u64 e = 0; e < E; e++) {
      orc_acc* a = &tot[g * E + e];
      a->sum = 0.0; a->count = 0; a->mn = 1.0e300; a->mx = -1.0e300;
    }
    for (orc_u64 t = 0; t < c.tasks; t++) {
      orc_u64 f = c.first_row[t * c.group_count + g];
      if (f < first[g]) first[g] = f;
      for (orc_u64 e = 0; e < E; e++) {
        orc_acc* a = &tot[g * E + e];
        orc_acc* b = &c.acc[(t * c.group_count + g) * E + e];
        a->sum += b->sum;
        a->count += b->count;
        if (b->mn < a->mn) a->mn = b->mn;
        if (b->mx > a->mx) a->mx = b->mx;
      }
    }
  }
  /* selection sort of matched groups by first fact row (group counts are
   * small in every join this oracle checks) */
  orc_u64 out = 0;
  for (;;) {
    orc_u64 best = ~(orc_u64)0, bg = 0;
    for (orc_u64 g = 0; g < c.group_count; g++)
      if (first[g] < best) { best = first[g]; bg = g; }
    if (best == ~(orc_u64)0) break;
    first[bg] = ~(orc_u64)0;
    if (out < max_groups) {
      for (orc_u32 k = 0; out_group_keys && k < group_count; k++)
        out_group_keys[out * group_count + k] = (orc_f64)group_value[4 * bg + k];
      for (orc_u64 e = 0; e < E; e++) out_res[out * E + e] = orc_finish(funcs[e], &tot[bg * E + e]);
    }
    out++;
  }
  free(first);
  free(tot);
  free(c.acc);
  free(c.first_row);
  free(c.heads);
  free(c.next);
  free(c.group_of_dim);
  free(group_value);
  return out;
}

/* qe_aggregate_check_sum_exact, query_exec.c:7028-7049 (called for every
 * result column, :7083): a SUM over a whole-number column (U32/I32/U64/I64)
 * whose value reached 2^53 fails the query.  Returns 1 when every such
 * SUM stays below 2^53 in magnitude. */
int orc_check_sum_exact(const orc_u32* funcs, const orc_u32* arg_types, orc_u32 expr_count, const orc_f64* res,
                        orc_u64 groups)
{
  for (orc_u32 e = 0; e < expr_count; e++) {
    const orc_u32 t = arg_types[e];
    if (funcs[e] != ORC_AGG_SUM || (t != ORC_TYPE_U32 && t != ORC_TYPE_I32 && t != ORC_TYPE_U64 && t != ORC_TYPE_I64)) continue;
    for (orc_u64 g = 0; g < groups; g++) {
      const orc_f64 v = res[g * expr_count + e];
      if (v >= 9007199254740992.0 || v <= -9007199254740992.0) return 0;
    }
  }
  return 1;
}

/* ------------------------------------------------------------------ */
/* numeric-AND count and dense group (CPU baselines at scale)           */
/* ------------------------------------------------------------------ */
typedef struct {
  const orc_column* cols;
  const orc_u32 *leaf_col, *leaf_op;
  const orc_f64* leaf_target;
  orc_u32 leaf_count;
  orc_u64 rows, tasks;
  orc_u64* counts;
} orc_and_ctx;

/* qe_cpu_scan_bits_task semantics (query_exec.c:1140-1186): a row passes
 * when every leaf's f64 comparison holds */
static void orc_and_task(void* raw, orc_u64 t)
{
  orc_and_ctx* c = (orc_and_ctx*)raw;
  orc_u64 lo, hi, k = 0;
  orc_divide_work(c->rows, c->tasks, t, &lo, &hi);
  for (orc_u64 i = lo; i < hi; i++) {
    int pass = 1;
    for (orc_u32 l = 0; l < c->leaf_count && pass; l++) {
      orc_f64 v = orc_load(&c->cols[c->leaf_col[l]], i), x = c->leaf_target[l];
      switch (c->leaf_op[l]) {
        case 0: pass = v == x; break;
        case 1: pass = v < x; break;
        case 2: pass = v <= x; break;
        case 3: pass = v > x; break;
        default: pass = v >= x; break;
      }
    }
    k += (orc_u64)pass;
  }
  c->counts[t] = k;
}

orc_u64 orc_scan_chunked_count(const orc_column* cols, const orc_u32* leaf_col, const orc_u32* leaf_op,
                               const orc_f64* leaf_target, orc_u32 leaf_count, orc_u64 rows, int workers)
{
  orc_and_ctx c = {cols, leaf_col, leaf_op, leaf_target, leaf_count, rows, 0, 0};
  if (workers < 1) workers = 1;
  c.tasks = rows / 65536 < (orc_u64)workers * 4 ? rows / 65536 : (orc_u64)workers * 4;
  if (c.tasks < 1) c.tasks = 1;
  c.counts = (orc_u64*)calloc(c.tasks, sizeof(orc_u64));
  orc_parallel_for(c.tasks, workers, orc_and_task, &c);
  orc_u64 total = 0;
  for (orc_u64 t = 0; t < c.tasks; t++) total += c.counts[t];
  free(c.counts);
  return total;
}

typedef struct {
  const orc_u32 *key, *arg;
  orc_u64 rows, tasks;
  orc_u32 key_min, range;
  orc_u64 *cnt, *sum;
  orc_u32 *mn, *mx;
} orc_dense_ctx;

static void orc_dense_task(void* raw, orc_u64 t)
{
  orc_dense_ctx* c = (orc_dense_ctx*)raw;
  orc_u64 lo, hi;
  orc_divide_work(c->rows, c->tasks, t, &lo, &hi);
  orc_u64* cnt = c->cnt + t * c->range;
  orc_u64* sum = c->sum + t * c->range;
  orc_u32* mn = c->mn + t * c->range;
  orc_u32* mx = c->mx + t * c->range;
  for (orc_u64 i = lo; i < hi; i++) {
    orc_u32 g = c->key[i] - c->key_min;
    orc_u32 v = c->arg ? c->arg[i] : 0;
    cnt[g] += 1;
    sum[g] += v;
    if (v < mn[g]) mn[g] = v;
    if (v > mx[g]) mx[g] = v;
  }
}

void orc_dense_group_u32(const orc_u32* key, const orc_u32* arg, orc_u64 rows, orc_u32 key_min,
                         orc_u32 range, int workers, orc_u64* out_count, orc_u64* out_sum,
                         orc_u32* out_min, orc_u32* out_max)
{
  orc_dense_ctx c;
  memset(&c, 0, sizeof(c));
  if (workers < 1) workers = 1;
  c.key = key; c.arg = arg; c.rows = rows; c.key_min = key_min; c.range = range;
  c.tasks = rows < 65536 ? 1 : (orc_u64)workers;
  c.cnt = (orc_u64*)calloc(c.tasks * range, sizeof(orc_u64));
  c.sum = (orc_u64*)calloc(c.tasks * range, sizeof(orc_u64));
  c.mn = (orc_u32*)malloc(sizeof(orc_u32) * c.tasks * range);
  c.mx = (orc_u32*)calloc(c.tasks * range, sizeof(orc_u32));
  for (orc_u64 i = 0; i < c.tasks * range; i++) c.mn[i] = 0xffffffffu;
  orc_parallel_for(c.tasks, workers, orc_dense_task, &c);
  for (orc_u32 g = 0; g < range; g++) {
    out_count[g] = 0; out_sum[g] = 0; out_min[g] = 0xffffffffu; out_max[g] = 0;
    for (orc_u64 t = 0; t < c.tasks; t++) {
      out_count[g] += c.cnt[t * range + g];
      out_sum[g] += c.sum[t * range + g];
      if (c.mn[t * range + g] < out_min[g]) out_min[g] = c.mn[t * range + g];
      if (c.mx[t * range + g] > out_max[g]) out_max[g] = c.mx[t * range + g];
    }
  }
  free(c.cnt);
  free(c.sum);
  free(c.mn);
  free(c.mx);
}

/* ------------------------------------------------------------------ */
/* hash join producing row pairs                                        */
/* ------------------------------------------------------------------ */
/* qe_hash_join_impl, query_exec.c:14541-15358 with hash_join_probe.comp:
 * numeric keys are compared as f64 values (the key columns are converted
 * to f64, :14704-14725), every build row is kept so duplicate build keys
 * fan out, a LEFT join pairs an unmatched probe row with all ones
 * (BUILD_NULL), SEMI / ANTI keep the probe positions with / without a
 * match, once each, in probe order (qe_join_filter_left_rows,
 * :14470-14520).  NaN keys match nothing (f64 ==), -0.0 matches 0.0.
 * Records are written in probe-position order and, within one probe row,
 * build-position order (the reference's own order is that of its atomics).
 * join_type: 0 inner, 1 left, 2 semi, 3 anti.  Returns the record count;
 * writes at most cap records (pairs for inner/left, positions otherwise). */
orc_u64 orc_hash_join(const orc_column* probe_key, const orc_u64* probe_rows, orc_u64 np,
                      const orc_column* build_key, const orc_u64* build_rows, orc_u64 nb, orc_u32 join_type,
                      orc_u64* out, orc_u64 cap)
{
  orc_u64 size = 16;
  while (size < 2 * nb) size <<= 1;
  /* distinct keys: first and last build position of each, chained through
   * next[] in ascending position order */
  orc_f64* skey = (orc_f64*)malloc(size * sizeof(orc_f64));
  orc_u64* first = (orc_u64*)malloc(size * sizeof(orc_u64));
  orc_u64* last = (orc_u64*)malloc(size * sizeof(orc_u64));
  orc_u64* next = (orc_u64*)malloc((nb ? nb : 1) * sizeof(orc_u64));
  for (orc_u64 i = 0; i < size; i++) first[i] = ~0ULL;
  for (orc_u64 i = 0; i < nb; i++) {
    orc_f64 v = orc_load(build_key, build_rows ? build_rows[i] : i);
    next[i] = ~0ULL;
    if (v != v) continue;
    orc_f64 kv = v;
    orc_u64 h = orc_hash_keys((const orc_f64* const[]){&kv}, 1, 0) & (size - 1);
    while (first[h] != ~0ULL && skey[h] != v) h = (h + 1) & (size - 1);
    if (first[h] == ~0ULL) {
      first[h] = last[h] = i;
      skey[h] = v;
    } else {
      next[last[h]] = i;
      last[h] = i;
    }
  }
  orc_u64 total = 0;
  for (orc_u64 p = 0; p < np; p++) {
    orc_f64 v = orc_load(probe_key, probe_rows ? probe_rows[p] : p);
    orc_u64 chain = ~0ULL;
    if (v == v) {
      orc_f64 kv = v;
      orc_u64 h = orc_hash_keys((const orc_f64* const[]){&kv}, 1, 0) & (size - 1);
      while (first[h] != ~0ULL && skey[h] != v) h = (h + 1) & (size - 1);
      chain = first[h];
    }
    if (join_type >= 2) {
      int keep = join_type == 2 ? chain != ~0ULL : chain == ~0ULL;
      if (keep) {
        if (total < cap) out[total] = p;
        total++;
      }
      continue;
    }
    if (chain == ~0ULL) {
      if (join_type == 1) {
        if (total < cap) {
          out[2 * total] = p;
          out[2 * total + 1] = ~0ULL;
        }
        total++;
      }
      continue;
    }
    for (orc_u64 b = chain; b != ~0ULL; b = next[b]) {
      if (total < cap) {
        out[2 * total] = p;
        out[2 * total + 1] = build_rows ? build_rows[b] : b;
      }
      total++;
    }
  }
  free(skey);
  free(first);
  free(last);
  free(next);
  return total;
}
