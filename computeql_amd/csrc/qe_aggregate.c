/*
 * qe_aggregate.c - global and grouped aggregates on the GPU.
 *
 *   global (no GROUP BY)   qe_aggregate_global_streamed :8842-9110 and
 *                          qe_aggregate_global_gpu :9309-9437 (whole columns);
 *                          qe_fused_filter_aggregate :9888-10253 (WHERE fused)
 *   dense GROUP BY         qe_dense_stream_core :7632-7980,
 *                          qe_aggregate_dense_streamed :7984-8049;
 *                          qe_fused_filter_group_aggregate :9442-9886 (WHERE fused)
 *   output typing          qe_aggregate_build_output :7053, qe_agg_func_policy
 *                          (query_exec.h:917-926)
 *
 * Arithmetic: U32/I32 arguments are summed as exact 64-bit integers and
 * converted to F64 once per group; MIN/MAX are exact and converted to F64.
 * The CPU path sums in F64 in row order, which is exact (and therefore
 * identical) while every partial sum stays below 2^53.  F32/F64/U64/I64
 * arguments are summed in F64: per thread over its rows in order, then a
 * fixed wave/block tree, then blocks in block order.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>

#include "qe_internal.h"

B32 qe_agg_plan(QE_Table* t, const U32* funcs, const S32* arg_cols, U32 expr_count, QE_AggPlan* p)
{
  memset(p, 0, sizeof(*p));
  if (expr_count == 0 || expr_count > QE_AGG_MAX_EXPRS) {
    qe_set_error("aggregate: 1..%d expressions", QE_AGG_MAX_EXPRS);
    return 0;
  }
  p->expr_count = expr_count;
  for (U32 e = 0; e < expr_count; e++) {
    U32 f = funcs[e];
    if (f > QE_AGG_FUNC_MAX) {
      qe_set_error("aggregate: unsupported function %u", f);
      return 0;
    }
    p->funcs[e] = f;
    p->expr_arg[e] = -1;
    if (arg_cols[e] < 0) {
      if (f != QE_AGG_FUNC_COUNT) {
        qe_set_error("aggregate: function %u needs an argument", f);
        return 0;
      }
      continue;
    }
    if ((U32)arg_cols[e] >= t->column_count) {
      qe_set_error("aggregate: unknown column %d", arg_cols[e]);
      return 0;
    }
    if (f == QE_AGG_FUNC_COUNT) continue; /* COUNT(col): no NULLs on this path */
    U32 a = 0;
    while (a < p->arg_count && p->arg_col[a] != (U32)arg_cols[e]) a++;
    if (a == p->arg_count) {
      if (p->arg_count == QE_AGG_MAX_ARGS) {
        qe_set_error("aggregate: more than %d distinct argument columns", QE_AGG_MAX_ARGS);
        return 0;
      }
      p->arg_col[p->arg_count++] = (U32)arg_cols[e];
    }
    p->expr_arg[e] = (S32)a;
    if (f == QE_AGG_FUNC_SUM || f == QE_AGG_FUNC_AVG) p->need_sum = 1;
    if (f == QE_AGG_FUNC_MIN || f == QE_AGG_FUNC_MAX) p->need_minmax = 1;
  }
  return 1;
}

static B32 qe_type_int32(U32 t) { return t == CQL_ColumnType_U32 || t == CQL_ColumnType_I32; }

static F64 qe_key_to_f64(U32 type, U64 key)
{
  switch (type) {
    case CQL_ColumnType_U32: return (F64)(U32)key;
    case CQL_ColumnType_I32: return (F64)(S32)((U32)key ^ 0x80000000u);
    case CQL_ColumnType_U64: return (F64)key;
    case CQL_ColumnType_I64: return (F64)(S64)(key ^ 0x8000000000000000ull);
    default: {
      U64 b = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
      F64 d;
      memcpy(&d, &b, 8);
      return d;
    }
  }
}

static F64 qe_sum_to_f64(U32 type, U64 sum)
{
  if (type == CQL_ColumnType_U32) return (F64)sum;
  if (type == CQL_ColumnType_I32) return (F64)(S64)sum;
  F64 d;
  memcpy(&d, &sum, 8);
  return d;
}

/* ---------------- global aggregate (one pass, optional fused WHERE) ---------------- */
/* the columns, arg slots / type bytes and leaf words aggregate_global takes */
typedef struct QE_GlobalInputs {
  QE_ColumnSlots slots;
  U32 slot_bits, type_bits;
  U32 leaf_words[QE_WHERE_WORDS];
  U32 leaf_bytes, leaf_packed;
} QE_GlobalInputs;

/* a WHERE's device words over a kernel's column slots (added to `slots`
 * as needed): the leaves with NULL bitmaps and encodings, then the terms */
static B32 qe_where_prepare(QE_Table* t, const QE_Where* w, QE_ColumnSlots* slots, B32 enc, U32* words, U32* bytes,
                            U32* packed)
{
  *bytes = *packed = 0;
  if (!w || !w->leaf_count) return 1;
  U32 slot[QE_SCAN_MAX_LEAVES];
  for (U32 l = 0; l < w->leaf_count; l++) {
    slot[l] = qe_slot_of(slots, w->leaf[l].column);
    if (slot[l] == ~0u) {
      qe_set_error("aggregate: too many columns");
      return 0;
    }
  }
  *bytes = qe_where_words(t, w, slot, 1, enc, words, packed);
  return 1;
}

static B32 qe_global_prepare(QE_Table* t, const QE_Where* w, const QE_AggPlan* p, QE_GlobalInputs* in)
{
  memset(in, 0, sizeof(*in));
  /* encoded columns (a3) are read as codes: an encoded arg is a U32 column
   * of codes (type byte U32 | encoding << 4) whose base is added back by
   * qe_global_decode */
  for (U32 a = 0; a < p->arg_count; a++) {
    const U32 slot = qe_slot_of(&in->slots, p->arg_col[a]);
    in->slot_bits |= slot << (4 * a);
    U32 enc = qe_col_enc(t, p->arg_col[a]);
    in->type_bits |= (enc ? (CQL_ColumnType_U32 | (enc << 4)) : t->columns[p->arg_col[a]].type) << (8 * a);
  }
  return qe_where_prepare(t, w, &in->slots, 1, in->leaf_words, &in->leaf_bytes, &in->leaf_packed);
}

/* aggregate_global_fold's 13 result words (count, then sum/min/max per
 * argument) into the query's cells */
static void qe_global_decode(QE_Table* t, const QE_AggPlan* p, U64* res, U64* out_cells)
{
  U64 count = res[0];
  /* encoded args: the kernel saw codes = value - base */
  for (U32 a = 0; a < p->arg_count && count; a++) {
    const QE_Column* col = &t->columns[p->arg_col[a]];
    if (!qe_col_enc(t, p->arg_col[a])) continue;
    const U32 base = col->enc_base;
    if (col->type == CQL_ColumnType_I32) {
      res[1 + 3 * a] = (U64)((S64)res[1 + 3 * a] + (S64)count * (S64)(S32)base);
      res[2 + 3 * a] = (U32)(base + (U32)res[2 + 3 * a]) ^ 0x80000000u;
      res[3 + 3 * a] = (U32)(base + (U32)res[3 + 3 * a]) ^ 0x80000000u;
    } else {
      res[1 + 3 * a] += count * (U64)base;
      res[2 + 3 * a] = (U32)(base + (U32)res[2 + 3 * a]);
      res[3 + 3 * a] = (U32)(base + (U32)res[3 + 3 * a]);
    }
  }
  for (U32 e = 0; e < p->expr_count; e++) {
    S32 a = p->expr_arg[e];
    F64 sum = 0.0, mn = 0.0, mx = 0.0;
    if (a >= 0 && count > 0) {
      U32 type = t->columns[p->arg_col[a]].type;
      sum = qe_sum_to_f64(type, res[1 + 3 * a]);
      mn = qe_key_to_f64(type, res[2 + 3 * a]);
      mx = qe_key_to_f64(type, res[3 + 3 * a]);
      if (qe_type_int32(type)) {
        out_cells[e] = qe_agg_cell_int(p->funcs[e], count, (S64)res[1 + 3 * a], mn, mx, qe_t_raw_int_sums);
        continue;
      }
    }
    out_cells[e] = qe_agg_cell(p->funcs[e], count, sum, mn, mx);
  }
}

static B32 qe_aggregate_global(QE_Table* t, const QE_Where* w, const QE_Prune* prune, const QE_AggPlan* p,
                               U64* out_cells, QE_Trace* trace)
{
  const B32 terms = w && w->term_count;
  GPU_Kernel* k = qe_kernel(terms ? "aggregate_global_or" : "aggregate_global");
  GPU_Kernel* kf = qe_kernel("aggregate_global_fold");
  if (!k || !kf) return 0;
  QE_GlobalInputs in;
  if (!qe_global_prepare(t, w, p, &in)) return 0;
  const U32 leaf_count = w ? w->leaf_count : 0;
  const QE_ColumnSlots slots = in.slots;
  U64 r0 = prune->row_begin, r1 = prune->row_end;
  U64 quads = (r1 + 3) / 4 - r0 / 4;
  U32 blocks = qe_stream_blocks(quads);
  GPU_Buffer* result = qe_scratch("agg_global_result", 16 * 8);
  GPU_Buffer* partials = qe_scratch("agg_global_partials", (U64)blocks * 13 * 8);
  GPU_Buffer* leaf_buf = leaf_count ? qe_scratch("agg_global_leaves", sizeof(in.leaf_words)) : 0;
  GPU_Buffer* skip_buf = prune->skip_bits ? qe_scratch("agg_global_skip", prune->skip_words * 4) : 0;
  B32 ok = result && partials && (!leaf_count || leaf_buf) && (!prune->skip_bits || skip_buf);
  U64 res[16];
  memset(res, 0, sizeof(res));
  if (ok) {
    if (leaf_buf) gpu_kernel_set_arg_buffer(k, 0, leaf_buf);
    else gpu_hip_kernel_clear_arg(k, 0);
    gpu_kernel_set_arg_buffer(k, 1, result);
    gpu_kernel_set_arg_buffer(k, 2, partials);
    if (skip_buf) gpu_kernel_set_arg_buffer(k, 5, skip_buf);
    else gpu_hip_kernel_clear_arg(k, 5);
    for (U32 s = 0; s < QE_MAX_COLUMN_BINDINGS; s++) {
      if (s < slots.count) gpu_kernel_set_arg_buffer(k, 6 + s, qe_col_kbuf(t, slots.column[s]));
      else gpu_hip_kernel_clear_arg(k, 6 + s);
    }
    gpu_kernel_set_arg_u64(k, 0, r1 - r0);
    gpu_kernel_set_arg_u64(k, 3, QE_ZONE_SHIFT);
    gpu_kernel_set_arg_u64(k, 4, skip_buf ? 1 : 0);
    gpu_kernel_set_arg_u64(k, 5, r0);
    gpu_kernel_set_arg_u64(k, 7, in.leaf_packed);
    qe_kernel_set_word(k, QK_W_ARGS, p->arg_count | (in.slot_bits << 8));
    qe_kernel_set_word(k, QK_W_TYPES, in.type_bits);
    /* per-block partials, then one block folds them in block order */
    gpu_kernel_set_arg_buffer(kf, 1, result);
    gpu_kernel_set_arg_buffer(kf, 2, partials);
    gpu_kernel_set_arg_u64(kf, 0, blocks);
    qe_kernel_set_word(kf, QK_W_ARGS, p->arg_count);
    qe_kernel_set_word(kf, QK_W_TYPES, in.type_bits);
    GPU_Batch* b = gpu_batch_begin(0, sizeof(res));
    if (leaf_buf) gpu_batch_buffer_write(b, leaf_buf, in.leaf_words, in.leaf_bytes);
    if (skip_buf) gpu_batch_buffer_write(b, skip_buf, prune->skip_bits, prune->skip_words * 4);
    gpu_batch_kernel_execute(b, k, blocks * QE_WG, QE_WG);
    gpu_batch_kernel_execute(b, kf, QE_WG, QE_WG);
    gpu_batch_buffer_read(b, result, res, sizeof(res));
    ok = gpu_batch_end(b);
  }
  qe_scratch_release(result);
  qe_scratch_release(partials);
  qe_scratch_release(leaf_buf);
  qe_scratch_release(skip_buf);
  if (!ok) {
    qe_set_error("aggregate_global: %s", gpu_hip_last_error());
    return 0;
  }
  qe_global_decode(t, p, res, out_cells);
  if (trace) {
    trace->path = leaf_count ? QE_AggPath_FusedGlobal : QE_AggPath_Global;
    trace->kernel_count = 2;
    trace->rows_scanned = r1 - r0;
    trace->rows_after = res[0];
    U64 bytes = 0;
    for (U32 s = 0; s < slots.count; s++) bytes += (r1 - r0) * t->columns[slots.column[s]].width;
    trace->bytes_read = bytes;
    snprintf(trace->kernel, sizeof(trace->kernel), "aggregate_global");
  }
  return 1;
}

/* ---------------- the same without allocating ---------------- */
/* Scratch layout (QE_NOALLOC_SCRATCH_BYTES): leaf words [0, 256) | result
 * words [256, 384) | per-block partials from 512, 104 bytes a block. */
#define QE_NOALLOC_LEAVES 0u
#define QE_NOALLOC_RESULT 256u
#define QE_NOALLOC_PARTIALS 512u
#define QE_NOALLOC_MAX_BLOCKS ((QE_NOALLOC_SCRATCH_BYTES - QE_NOALLOC_PARTIALS) / (13u * 8u))

static void qe_args_bind(QK_Args* a, U32 index, GPU_Buffer* b)
{
  a->ptr[index] = b ? gpu_hip_buffer_device_address(b) : 0;
  a->bytes[index] = b ? gpu_hip_buffer_size(b) : 0;
}

B32 qe_aggregate_noalloc(QE_Table* table, const QE_IRNode* where, U32 where_count, const U32* funcs,
                         const S32* arg_cols, U32 expr_count, U64* out_cells, QE_Trace* trace)
{
  U64 t0, p0;
  qe_trace_begin(trace, &t0, &p0);
  if (!table || !funcs || !arg_cols || !out_cells) {
    qe_set_error("qe_aggregate_noalloc: null argument");
    return 0;
  }
  if (table->host_backed) {
    qe_set_error("qe_aggregate_noalloc: the table's columns are not in device memory");
    return 0;
  }
  if (!qe_where_non_null(table, where, where_count, "qe_aggregate_noalloc")) return 0;
  for (U32 e = 0; e < expr_count; e++)
    if (!qe_require_non_null(table, arg_cols[e], "qe_aggregate_noalloc")) return 0;
  QE_AggPlan plan;
  if (!qe_agg_plan(table, funcs, arg_cols, expr_count, &plan)) return 0;
  QE_Where w;
  memset(&w, 0, sizeof(w));
  if (where && where_count) {
    static __thread QE_Bytecode prog;
    if (!qe_bytecode_program_build(table, where, where_count, &prog)) {
      qe_set_error("qe_aggregate_noalloc: %s", prog.error);
      return 0;
    }
    if (!qe_where_build(table, where, where_count, &w)) {
      qe_set_error("qe_aggregate_noalloc: the WHERE must be an AND / OR of at most %d column-literal comparisons "
                   "(%d conjunctions)", QE_SCAN_MAX_LEAVES, QE_WHERE_MAX_TERMS);
      return 0;
    }
  }
  const U32 leaf_count = w.leaf_count;
  const U64 scratch = gpu_hip_noalloc_scratch();
  if (!scratch) {
    qe_set_error("qe_aggregate_noalloc: %s", *gpu_hip_last_error() ? gpu_hip_last_error() : "no scratch on this device");
    return 0;
  }
  QE_GlobalInputs in;
  if (!qe_global_prepare(table, &w, &plan, &in)) return 0;
  const U64 rows = table->row_count;
  U64 res[16];
  memset(res, 0, sizeof(res));
  U32 blocks = qe_stream_blocks((rows + 3) / 4);
  if (blocks > QE_NOALLOC_MAX_BLOCKS) blocks = QE_NOALLOC_MAX_BLOCKS; /* rows are grid-strided */
  if (rows) {
    QK_Args a[2];
    memset(a, 0, sizeof(a));
    QK_Args* k = &a[0];
    if (leaf_count) {
      k->ptr[0] = scratch + QE_NOALLOC_LEAVES;
      k->bytes[0] = in.leaf_bytes;
    }
    k->ptr[1] = scratch + QE_NOALLOC_RESULT;
    k->bytes[1] = 128;
    k->ptr[2] = scratch + QE_NOALLOC_PARTIALS;
    k->bytes[2] = (U64)blocks * 13 * 8;
    for (U32 s = 0; s < in.slots.count; s++) qe_args_bind(k, 6 + s, qe_col_kbuf(table, in.slots.column[s]));
    k->words[0] = (U32)rows;
    k->words[1] = (U32)(rows >> 32);
    k->words[6] = QE_ZONE_SHIFT;
    k->words[14] = in.leaf_packed;
    k->words[QK_W_ARGS] = plan.arg_count | (in.slot_bits << 8);
    k->words[QK_W_TYPES] = in.type_bits;
    k->global_work_size = blocks * QE_WG;
    k->grid_blocks = blocks;
    QK_Args* f = &a[1];
    f->ptr[1] = scratch + QE_NOALLOC_RESULT;
    f->bytes[1] = 128;
    f->ptr[2] = scratch + QE_NOALLOC_PARTIALS;
    f->bytes[2] = (U64)blocks * 13 * 8;
    f->words[0] = blocks;
    f->words[QK_W_ARGS] = plan.arg_count;
    f->words[QK_W_TYPES] = in.type_bits;
    f->global_work_size = QE_WG;
    f->grid_blocks = 1;
    const char* names[2] = {w.term_count ? "aggregate_global_or" : "aggregate_global", "aggregate_global_fold"};
    const U32 grids[2] = {blocks, 1};
    if (!gpu_hip_noalloc_run(in.leaf_words, in.leaf_bytes, QE_NOALLOC_LEAVES, names, a, grids, 2, res,
                             13 * 8, QE_NOALLOC_RESULT)) {
      qe_set_error("qe_aggregate_noalloc: %s", gpu_hip_last_error());
      return 0;
    }
  }
  qe_global_decode(table, &plan, res, out_cells);
  if (!qe_check_sum_exact(table, funcs, arg_cols, expr_count, out_cells, 1)) return 0;
  if (trace) {
    trace->path = leaf_count ? QE_AggPath_FusedGlobal : QE_AggPath_Global;
    trace->kernel_count = rows ? 2 : 0;
    trace->rows_before = trace->rows_scanned = rows;
    trace->rows_after = res[0];
    U64 bytes = 0;
    for (U32 s = 0; s < in.slots.count; s++) bytes += rows * table->columns[in.slots.column[s]].width;
    trace->bytes_read = bytes;
    snprintf(trace->kernel, sizeof(trace->kernel), "aggregate_global");
  }
  qe_trace_end(trace, t0, p0);
  return 1;
}

/* ---------------- GROUP BY keys ---------------- */
B32 qe_group_keys(const QE_Table* t, const S32* cols, U32 count, QE_GroupKeys* out)
{
  memset(out, 0, sizeof(*out));
  if (count > QE_AGG_MAX_GROUP_COLS || (count && !cols)) {
    qe_set_error("aggregate: at most %d GROUP BY columns", QE_AGG_MAX_GROUP_COLS);
    return 0;
  }
  out->count = count;
  for (U32 k = 0; k < count; k++) {
    if (cols[k] < 0 || (U32)cols[k] >= t->column_count) {
      qe_set_error("aggregate: unknown group column %d", cols[k]);
      return 0;
    }
    const QE_Column* c = &t->columns[cols[k]];
    out->col[k] = (U32)cols[k];
    out->type[k] = c->type;
    if (qe_type_int32(c->type) && c->zonemap_valid && c->max_value >= c->min_value) {
      out->min_code[k] = c->type == CQL_ColumnType_I32 ? ((U32)(S32)c->min_value ^ 0x80000000u) : (U32)c->min_value;
      out->range[k] = (U64)(c->max_value - c->min_value) + 1;
    }
  }
  return 1;
}

U64 qe_group_keys_span(const QE_GroupKeys* k)
{
  U64 span = 1;
  for (U32 i = 0; i < k->count; i++) {
    if (!k->range[i]) return 0;
    if (span > ~0ull / k->range[i]) return ~0ull;
    span *= k->range[i];
  }
  return span;
}

F64 qe_group_key_value(const QE_GroupKeys* k, U32 i, U64 d)
{
  U32 code = k->min_code[i] + (U32)d;
  return k->type[i] == CQL_ColumnType_I32 ? (F64)(S32)(code ^ 0x80000000u) : (F64)code;
}

/* the keys' values of combined group number g (digits, key 0 most
 * significant) into out[0..K) */
static void qe_group_digits(const QE_GroupKeys* k, U64 g, F64* out)
{
  for (U32 i = k->count; i-- > 0;) {
    U64 d = g % k->range[i];
    g /= k->range[i];
    if (out) out[i] = qe_group_key_value(k, i, d);
  }
}

/* ---------------- dense GROUP BY ---------------- */
/* per-group device totals of aggregate_dense / join_probe: count u64[R];
 * sum u64[E][R]; min u32[E][R]; max u32[E][R] */

static B32 qe_dense_totals_alloc(QE_DenseTotals* d, U32 range, U32 args, B32 wide)
{
  d->range = range;
  d->args = args;
  d->wide = wide;
  U64 bytes = (U64)range * 8 * (1 + args) + (U64)range * (wide ? 8 : 4) * 2 * args;
  d->block = malloc(bytes ? bytes : 8);
  if (!d->block) return 0;
  d->count = (U64*)d->block;
  d->sum = d->count + range;
  d->mn = (U32*)(d->sum + (U64)range * args);
  d->mx = d->mn + (U64)range * args;
  d->mn64 = d->sum + (U64)range * args;
  d->mx64 = d->mn64 + (U64)range * args;
  return 1;
}

U64 qe_dense_totals_bytes(U32 range, U32 args) { return (U64)range * 8 * (1 + args) + (U64)range * 4 * 2 * args; }

/* LDS bytes of a dense GROUP BY block: count u32 | sums u64 | min/max u32
 * per argument (u64 when an argument is wider than a 32-bit integer);
 * narrow tables count+carries u32 | sum low word u32 | min/max u32 for
 * one argument */
U64 qe_dense_lds(U32 range, U32 args, B32 need_sum, B32 need_minmax, B32 narrow)
{
  if (narrow) return (U64)range * (4 + (need_sum ? 4 : 0) + (need_minmax ? 8 : 0));
  return (U64)range * 4 + (need_sum ? (U64)args * range * 8 : 0) + (need_minmax ? (U64)args * range * 8 : 0);
}

static U64 qe_dense_lds_wide(U32 range, U32 args, B32 need_sum, B32 need_minmax)
{
  return (U64)range * 4 + (need_sum ? (U64)args * range * 8 : 0) + (need_minmax ? (U64)args * range * 16 : 0);
}

/* arguments other than 32-bit integers take the wide tables */
static B32 qe_dense_wide_args(const QE_Table* t, const U32* arg_cols, U32 args)
{
  for (U32 a = 0; a < args; a++)
    if (!qe_type_int32(t->columns[arg_cols[a]].type)) return 1;
  return 0;
}

/* the narrow kernels take one U32 argument (or its narrow codes) and at
 * least one of SUM and MIN/MAX */
B32 qe_dense_narrow_ok(const QE_Table* t, const U32* arg_cols, U32 args, B32 need_sum, B32 need_minmax, U32 range)
{
  return qe_settings_u64("QE_AGG_DENSE_NARROW", 1) != 0 && args == 1 && (need_sum || need_minmax) &&
         t->columns[arg_cols[0]].type == CQL_ColumnType_U32 &&
         qe_dense_lds(range, 1, need_sum, need_minmax, 1) <= gpu_hip_max_lds_bytes();
}

/* runs aggregate_dense over the keys of `keys` (columns of `t`) and args;
 * fills totals indexed by the keys' combined digits */
B32 qe_dense_run(QE_Table* t, const QE_GroupKeys* keys, const U32* arg_cols, U32 arg_count, B32 need_sum,
                 B32 need_minmax, const QE_Where* w, const QE_Prune* prune, QE_DenseTotals* out,
                 QE_Trace* trace)
{
  const B32 terms = w && w->term_count;
  /* arguments wider than 32-bit integers: the wide tables, over full-width
   * columns (no encodings) */
  const B32 wide = qe_dense_wide_args(t, arg_cols, arg_count);
  GPU_Kernel* k = qe_kernel(wide ? (terms ? "aggregate_dense_wide_or" : "aggregate_dense_wide")
                                 : (terms ? "aggregate_dense_or" : "aggregate_dense"));
  if (!k) return 0;
  const U64 span = qe_group_keys_span(keys);
  if (keys->count < 1 || keys->count > QE_DENSE_MAX_KEYS || span == 0 || span > 0xffffffffull) {
    qe_set_error("dense group by: 1..%d 32-bit integer keys with known ranges", QE_DENSE_MAX_KEYS);
    return 0;
  }
  const U32 range = (U32)span;
  U32 flags = (need_sum ? 1u : 0u) | (need_minmax ? 2u : 0u);
  U64 lds = wide ? qe_dense_lds_wide(range, arg_count, need_sum, need_minmax)
                 : qe_dense_lds(range, arg_count, need_sum, need_minmax, 0);
  U32 max_lds = gpu_hip_max_lds_bytes();
  /* the narrow tables (aggregate_dense_narrow_*) when the full ones do not
   * fit or hold a CU to one 256-thread block */
  const B32 narrow = !wide && qe_dense_narrow_ok(t, arg_cols, arg_count, need_sum, need_minmax, range) &&
                     (lds > max_lds || lds > qe_settings_u64("QE_AGG_DENSE_NARROW_ABOVE", 65536));
  if (narrow) lds = qe_dense_lds(range, arg_count, need_sum, need_minmax, 1);
  if (lds > max_lds) {
    qe_set_error("dense group by: %llu bytes of LDS needed, %u available", (unsigned long long)lds, max_lds);
    return 0;
  }
  if (narrow) {
    const char* nk = need_sum && need_minmax ? "aggregate_dense_narrow_sm"
                                             : (need_sum ? "aggregate_dense_narrow_s" : "aggregate_dense_narrow_m");
    char name[64];
    snprintf(name, sizeof(name), "%s%s", nk, terms ? "_or" : "");
    k = qe_kernel(name);
    if (!k) return 0;
  }
  const U32 wg = narrow ? 1024u : QE_WG;
  QE_ColumnSlots slots;
  memset(&slots, 0, sizeof(slots));
  /* encoded columns (a3) are read as codes: type byte U32 | encoding << 4,
   * the keys' minima rewritten to codes, the base added back below */
  U32 kslot[QE_DENSE_MAX_KEYS], kbyte[QE_DENSE_MAX_KEYS], kmin[QE_DENSE_MAX_KEYS];
  for (U32 i = 0; i < keys->count; i++) {
    const U32 col = keys->col[i];
    kslot[i] = qe_slot_of(&slots, col);
    const U32 enc = wide ? 0u : qe_col_enc(t, col);
    kbyte[i] = t->columns[col].type;
    kmin[i] = keys->min_code[i];
    if (enc) {
      kbyte[i] = CQL_ColumnType_U32 | (enc << 4);
      kmin[i] = keys->min_code[i] - t->columns[col].enc_base - (t->columns[col].type == CQL_ColumnType_I32 ? 0x80000000u : 0u);
    }
  }
  U32 slot_bits = 0, type_bits = 0;
  for (U32 a = 0; a < arg_count; a++) {
    U32 s = qe_slot_of(&slots, arg_cols[a]);
    U32 ty = t->columns[arg_cols[a]].type;
    U32 enc = wide ? 0u : qe_col_enc(t, arg_cols[a]);
    slot_bits |= s << (4 * a);
    type_bits |= (enc ? (CQL_ColumnType_U32 | (enc << 4)) : ty) << (8 * a);
  }
  U32 leaf_words[QE_WHERE_WORDS], leaf_bytes = 0, leaf_packed = 0;
  if (!qe_where_prepare(t, w, &slots, !wide, leaf_words, &leaf_bytes, &leaf_packed)) return 0;
  const U32 leaf_count = w ? w->leaf_count : 0;
  for (U32 i = 0; i < keys->count; i++)
    if (kslot[i] == ~0u) {
      qe_set_error("dense group by: too many columns");
      return 0;
    }
  /* several keys: their descriptors (u32 count, then slot | type << 8,
   * code minimum, range per key) in binding 2 */
  U32 kdesc[1 + 3 * QE_DENSE_MAX_KEYS];
  kdesc[0] = keys->count;
  for (U32 i = 0; i < keys->count; i++) {
    kdesc[1 + 3 * i] = kslot[i] | (kbyte[i] << 8);
    kdesc[2 + 3 * i] = kmin[i];
    kdesc[3 + 3 * i] = (U32)keys->range[i];
  }
  if (!qe_dense_totals_alloc(out, range, arg_count, wide)) return 0;
  U64 r0 = prune->row_begin, r1 = prune->row_end;
  U64 quads = (r1 + 3) / 4 - r0 / 4;
  /* one LDS table per block: fewer, fuller blocks when the table is big */
  U32 blocks = qe_stream_blocks(quads);
  U64 cus = gpu_hip_compute_units();
  U64 lds_blocks = lds ? (160ull * 1024 / (lds + 1024)) : 8;
  if (lds_blocks < 1) lds_blocks = 1;
  if (lds_blocks > qe_setting_blocks_per_cu()) lds_blocks = qe_setting_blocks_per_cu();
  if (narrow) {
    /* 1024-thread blocks of 4096-quad steps: as many as fit, one step each at least */
    U64 need = (quads + 4096 - 1) / 4096;
    blocks = (U32)(need < cus * lds_blocks ? (need ? need : 1) : cus * lds_blocks);
  }
  if ((U64)blocks > cus * lds_blocks) blocks = (U32)(cus * lds_blocks);
  const U64 mm_width = wide ? 8 : 4;
  U64 gbytes = (U64)range * 8 * (1 + arg_count) + (U64)range * mm_width * 2 * arg_count;
  GPU_Buffer* acc = qe_scratch("dense_totals", gbytes);
  GPU_Buffer* leaf_buf = leaf_count ? qe_scratch("dense_leaves", sizeof(leaf_words)) : 0;
  GPU_Buffer* skip_buf = prune->skip_bits ? qe_scratch("dense_skip", prune->skip_words * 4) : 0;
  GPU_Buffer* key_buf = keys->count > 1 ? qe_scratch("dense_keys", sizeof(kdesc)) : 0;
  B32 ok = acc && (!leaf_count || leaf_buf) && (!prune->skip_bits || skip_buf) && (keys->count == 1 || key_buf);
  if (ok) {
    if (leaf_buf) gpu_kernel_set_arg_buffer(k, 0, leaf_buf);
    else gpu_hip_kernel_clear_arg(k, 0);
    gpu_kernel_set_arg_buffer(k, 1, acc);
    if (key_buf) gpu_kernel_set_arg_buffer(k, 2, key_buf);
    else gpu_hip_kernel_clear_arg(k, 2);
    if (skip_buf) gpu_kernel_set_arg_buffer(k, 5, skip_buf);
    else gpu_hip_kernel_clear_arg(k, 5);
    for (U32 s = 0; s < QE_MAX_COLUMN_BINDINGS; s++) {
      if (s < slots.count)
        gpu_kernel_set_arg_buffer(k, 6 + s, wide ? t->columns[slots.column[s]].buffer : qe_col_kbuf(t, slots.column[s]));
      else gpu_hip_kernel_clear_arg(k, 6 + s);
    }
    gpu_kernel_set_arg_u64(k, 0, r1 - r0);
    gpu_kernel_set_arg_u64(k, 3, QE_ZONE_SHIFT);
    gpu_kernel_set_arg_u64(k, 4, skip_buf ? 1 : 0);
    gpu_kernel_set_arg_u64(k, 5, r0);
    gpu_kernel_set_arg_u64(k, 7, leaf_packed);
    qe_kernel_set_word(k, QK_W_KEY_MIN, kmin[0]);
    qe_kernel_set_word(k, QK_W_RANGE, range);
    qe_kernel_set_word(k, QK_W_ARGS, arg_count | (slot_bits << 8));
    qe_kernel_set_word(k, QK_W_TYPES, type_bits);
    qe_kernel_set_word(k, QK_W_KEY, kslot[0] | (kbyte[0] << 8) | (flags << 16));
    gpu_hip_kernel_set_lds_bytes(k, (U32)lds);
    GPU_Batch* b = gpu_batch_begin(0, gbytes);
    if (leaf_buf) gpu_batch_buffer_write(b, leaf_buf, leaf_words, leaf_bytes);
    if (skip_buf) gpu_batch_buffer_write(b, skip_buf, prune->skip_bits, prune->skip_words * 4);
    if (key_buf) gpu_batch_buffer_write(b, key_buf, kdesc, 4 * (1 + 3 * (U64)keys->count));
    /* count/sum/max start at 0, min at all ones: fills run from offset 0,
     * so zero all, fill through the min array, zero count+sum again */
    U64 zero_bytes = (U64)range * 8 * (1 + arg_count);
    gpu_batch_buffer_zero(b, acc, gbytes);
    if (arg_count && need_minmax) {
      gpu_batch_buffer_fill(b, acc, zero_bytes + (U64)range * mm_width * arg_count, 0xffffffffu);
      gpu_batch_buffer_zero(b, acc, zero_bytes);
    }
    gpu_batch_kernel_execute(b, k, blocks * wg, wg);
    gpu_batch_buffer_read(b, acc, out->block, gbytes);
    ok = gpu_batch_end(b);
    gpu_hip_kernel_set_lds_bytes(k, 0);
    gpu_hip_kernel_clear_arg(k, 2);
  }
  qe_scratch_release(acc);
  qe_scratch_release(leaf_buf);
  qe_scratch_release(skip_buf);
  qe_scratch_release(key_buf);
  if (!ok) {
    free(out->block);
    out->block = 0;
    qe_set_error("aggregate_dense: %s", gpu_hip_last_error());
    return 0;
  }
  /* encoded args: per group, the kernel saw codes = value - base */
  for (U32 a = 0; a < arg_count && !wide; a++) {
    if (!qe_col_enc(t, arg_cols[a])) continue;
    const QE_Column* col = &t->columns[arg_cols[a]];
    const U32 base = col->enc_base, flip = col->type == CQL_ColumnType_I32 ? 0x80000000u : 0u;
    for (U32 g = 0; g < range; g++) {
      const U64 cnt = out->count[g];
      if (!cnt) continue;
      const U64 i = (U64)a * range + g;
      if (flip) out->sum[i] = (U64)((S64)out->sum[i] + (S64)cnt * (S64)(S32)base);
      else out->sum[i] += cnt * (U64)base;
      out->mn[i] = (base + out->mn[i]) ^ flip;
      out->mx[i] = (base + out->mx[i]) ^ flip;
    }
  }
  if (trace) {
    trace->kernel_count = 1;
    trace->rows_scanned = r1 - r0;
    U64 bytes = 0;
    for (U32 s = 0; s < slots.count; s++) bytes += (r1 - r0) * t->columns[slots.column[s]].width;
    trace->bytes_read = bytes;
    snprintf(trace->kernel, sizeof(trace->kernel), "%s",
             narrow ? "aggregate_dense_narrow" : (wide ? "aggregate_dense_wide" : "aggregate_dense"));
  }
  return 1;
}

void qe_dense_totals_free(QE_DenseTotals* d)
{
  free(d->block);
  d->block = 0;
}

/* ---------------- hash GROUP BY (query_exec.c:10258-12300) ---------------- */
static F64 qe_hash_key_to_f64(U32 type, U64 code)
{
  if (type == CQL_ColumnType_U32 || type == CQL_ColumnType_U64) return (F64)code;
  if (type == CQL_ColumnType_I32 || type == CQL_ColumnType_I64) return (F64)(S64)code;
  F64 d;
  memcpy(&d, &code, 8);
  return d;
}

/* order-preserving code of an f64 (the kernels' qk_f64_key) back to f64 */
static F64 qe_f64_from_order_code(U64 key)
{
  U64 b = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
  F64 d;
  memcpy(&d, &b, 8);
  return d;
}

/* how aggregate_hash forms a row's group code from the keys: 0 one key (its
 * own code), 1 digits of 32-bit integer keys, 2 a hash of the key tuple */
static U32 qe_hash_key_mode(const QE_GroupKeys* keys)
{
  if (keys->count <= 1) return 0;
  const U64 span = qe_group_keys_span(keys);
  return (span != 0 && span < (1ull << 63)) ? 1u : 2u;
}

/* GROUP BY over the global hash table of aggregate_hash (no keys: one
 * group).  Rows: the fused leaves over the pruned range, or `list` (device
 * row ids) when given.  Groups come out in table-slot order. */
static B32 qe_aggregate_hash(QE_Table* t, const QE_GroupKeys* keys, const QE_Where* w, const QE_Prune* prune, GPU_Buffer* list, U64 list_count, const QE_AggPlan* p, U64 cap,
                             U64* out_group_count, F64* out_keys, U64* out_cells, QE_Trace* trace)
{
  GPU_Kernel* k = qe_kernel(w && w->term_count ? "aggregate_hash_or" : "aggregate_hash");
  GPU_Kernel* kc = qe_kernel("hash_compact");
  if (!k || !kc) return 0;
  const U32 E = p->arg_count;
  const U32 K = keys->count;
  const U32 mode = qe_hash_key_mode(keys);
  const U32 T = mode == 2 ? K : 0u;
  QE_ColumnSlots slots;
  memset(&slots, 0, sizeof(slots));
  U32 kslot[QE_AGG_MAX_GROUP_COLS];
  for (U32 i = 0; i < K; i++) kslot[i] = qe_slot_of(&slots, keys->col[i]);
  U32 key_slot = K ? kslot[0] : 0;
  U32 key_type = K ? keys->type[0] : CQL_ColumnType_U32;
  U32 slot_bits = 0, type_bits = 0;
  for (U32 a = 0; a < E; a++) {
    slot_bits |= qe_slot_of(&slots, p->arg_col[a]) << (4 * a);
    type_bits |= t->columns[p->arg_col[a]].type << (8 * a);
  }
  /* the hash kernel reads full-width columns: no encodings */
  U32 leaf_words[QE_WHERE_WORDS], leaf_bytes = 0, leaf_packed = 0;
  if (!qe_where_prepare(t, w, &slots, 0, leaf_words, &leaf_bytes, &leaf_packed)) return 0;
  const U32 leaf_count = w ? w->leaf_count : 0;
  B32 slots_ok = slots.count <= QE_MAX_COLUMN_BINDINGS;
  for (U32 i = 0; i < K; i++) slots_ok = slots_ok && kslot[i] != ~0u;
  if (!slots_ok) {
    qe_set_error("hash group by: too many columns");
    return 0;
  }
  /* several keys: mode, K, per key slot | type << 8, code minimum, range,
   * then the tuple hash's seed */
  U32 kdesc[3 + 3 * QE_AGG_MAX_GROUP_COLS];
  kdesc[0] = mode;
  kdesc[1] = K;
  for (U32 i = 0; i < K; i++) {
    kdesc[2 + 3 * i] = kslot[i] | (keys->type[i] << 8);
    kdesc[3 + 3 * i] = keys->min_code[i];
    kdesc[4 + 3 * i] = (U32)keys->range[i];
  }
  U64 r0 = prune->row_begin, r1 = prune->row_end;
  U64 rows = list ? list_count : r1 - r0;
  /* table size: twice the groups there can be (rows, or the keys' value
   * span from the zone maps), grown and rerun if it fills */
  U64 est = rows ? rows : 1;
  if (K == 0) est = 1;
  else {
    U64 span = qe_group_keys_span(keys);
    if (span && span < est) est = span;
  }
  U32 log2 = 10;
  while (log2 < 28 && (1ull << log2) < 2 * est) log2++;
  U64 forced = qe_settings_u64("QE_AGG_HASH_LOG2", 0); /* initial size override */
  if (forced >= 4 && forced <= 28) log2 = (U32)forced;
  U32 L = (U32)qe_settings_u64("QE_AGG_HASH_LDS_SLOTS", 512);
  if (L & (L - 1)) L = 512;
  if (L > 4096) L = 4096;
  if (mode == 2) L = 0; /* tuple keys skip the LDS stage */
  U64 lds = (U64)L * (8 + 24 * (U64)E + 4);
  U32 flags = (p->need_sum ? 1u : 0u) | (p->need_minmax ? 2u : 0u) | (K == 0 ? 4u : 0u);
  U64 cus = gpu_hip_compute_units();
  U32 blocks;
  if (list) {
    U64 need = (list_count + QE_WG - 1) / QE_WG;
    blocks = (U32)(need < cus * 4 ? (need ? need : 1) : cus * 4);
  } else {
    blocks = qe_stream_blocks((r1 + 3) / 4 - r0 / 4);
  }
  U32 rec_words = 2 + 3 * E + 2 * T;
  GPU_Buffer* leaf_buf = leaf_count ? qe_scratch("hash_leaves", sizeof(leaf_words)) : 0;
  GPU_Buffer* skip_buf = (!list && prune->skip_bits) ? qe_scratch("hash_skip", prune->skip_words * 4) : 0;
  GPU_Buffer* flag_buf = qe_scratch("hash_flags", 64);
  GPU_Buffer* key_buf = mode ? qe_scratch("hash_keys", sizeof(kdesc)) : 0;
  B32 ok = (!leaf_count || leaf_buf) && (list || !prune->skip_bits || skip_buf) && flag_buf && (!mode || key_buf);
  U64 groups = 0;
  U64* recs = 0;
  B32 err_set = 0;
  U32 seed = 0;
  for (U32 attempt = 0; ok && attempt < 6; attempt++) {
    U64 C = (1ull << log2) + 1;
    U64 table_bytes = C * 8 * (2 + 3 * (U64)E + 2 * (U64)T);
    U32 cblocks = (U32)((C + 4095) / 4096);
    if (cblocks > 1024) cblocks = 1024;
    GPU_Buffer* table = qe_scratch("hash_table", table_bytes);
    GPU_Buffer* bc = qe_scratch("hash_block_counts", (U64)cblocks * 8);
    GPU_Buffer* out = qe_scratch("hash_groups", C * 8 * rec_words);
    if (!table || !bc || !out) {
      qe_scratch_release(table);
      qe_scratch_release(bc);
      qe_scratch_release(out);
      ok = 0;
      break;
    }
    kdesc[2 + 3 * K] = seed;
    if (leaf_buf) gpu_kernel_set_arg_buffer(k, 0, leaf_buf);
    else gpu_hip_kernel_clear_arg(k, 0);
    gpu_kernel_set_arg_buffer(k, 1, table);
    gpu_kernel_set_arg_buffer(k, 2, flag_buf);
    if (list) gpu_kernel_set_arg_buffer(k, 3, list);
    else gpu_hip_kernel_clear_arg(k, 3);
    if (key_buf) gpu_kernel_set_arg_buffer(k, 4, key_buf);
    else gpu_hip_kernel_clear_arg(k, 4);
    if (skip_buf) gpu_kernel_set_arg_buffer(k, 5, skip_buf);
    else gpu_hip_kernel_clear_arg(k, 5);
    for (U32 sl = 0; sl < QE_MAX_COLUMN_BINDINGS; sl++) {
      if (sl < slots.count) gpu_kernel_set_arg_buffer(k, 6 + sl, t->columns[slots.column[sl]].buffer);
      else gpu_hip_kernel_clear_arg(k, 6 + sl);
    }
    gpu_kernel_set_arg_u64(k, 0, rows);
    gpu_kernel_set_arg_u64(k, 3, QE_ZONE_SHIFT);
    gpu_kernel_set_arg_u64(k, 4, skip_buf ? 1 : 0);
    gpu_kernel_set_arg_u64(k, 5, list ? 0 : r0);
    gpu_kernel_set_arg_u64(k, 7, list ? 0 : leaf_packed);
    qe_kernel_set_word(k, QK_W_ARGS, E | (slot_bits << 8));
    qe_kernel_set_word(k, QK_W_TYPES, type_bits);
    qe_kernel_set_word(k, QK_W_KEY, key_slot | (key_type << 8) | (flags << 16));
    qe_kernel_set_word(k, QK_W_LOG2, log2);
    qe_kernel_set_word(k, QK_W_RANGE, L);
    gpu_hip_kernel_set_lds_bytes(k, (U32)lds);
    gpu_kernel_set_arg_buffer(kc, 1, table);
    gpu_kernel_set_arg_buffer(kc, 2, bc);
    gpu_kernel_set_arg_buffer(kc, 3, out);
    qe_kernel_set_word(kc, QK_W_ARGS, E | (T << 8));
    qe_kernel_set_word(kc, QK_W_LOG2, log2);
    U32 fl[2] = {0, 0};
    U64* bcounts = (U64*)malloc((U64)cblocks * 8);
    if (!bcounts) {
      ok = 0;
    } else {
      GPU_Batch* b = gpu_batch_begin(0, (U64)cblocks * 8 + 8);
      if (leaf_buf) gpu_batch_buffer_write(b, leaf_buf, leaf_words, leaf_bytes);
      if (skip_buf) gpu_batch_buffer_write(b, skip_buf, prune->skip_bits, prune->skip_words * 4);
      if (key_buf) gpu_batch_buffer_write(b, key_buf, kdesc, 4 * (3 + 3 * (U64)K));
      gpu_batch_buffer_zero(b, flag_buf, 8);
      /* keys, mins and key minima start at all ones: zero everything, then
       * fill the leading key + min + key-min arrays */
      gpu_batch_buffer_zero(b, table, table_bytes);
      gpu_batch_buffer_fill(b, table, C * 8 * (1 + (U64)E + (U64)T), 0xffffffffu);
      gpu_batch_kernel_execute(b, k, blocks * QE_WG, QE_WG);
      gpu_kernel_set_arg_u64(kc, 6, 1); /* push slot 6: mode, block counts */
      gpu_batch_kernel_execute(b, kc, cblocks * QE_WG, QE_WG);
      gpu_kernel_set_arg_u64(kc, 6, 2); /* write records */
      gpu_batch_kernel_execute(b, kc, cblocks * QE_WG, QE_WG);
      gpu_batch_buffer_read(b, flag_buf, fl, 8);
      gpu_batch_buffer_read(b, bc, bcounts, (U64)cblocks * 8);
      ok = gpu_batch_end(b);
      gpu_hip_kernel_set_lds_bytes(k, 0);
      gpu_hip_kernel_clear_arg(k, 4);
    }
    B32 full = ok && fl[0] != 0;
    if (ok && fl[1]) {
      qe_set_error("hash group by: a key lies outside its column's zone-map range");
      err_set = 1;
      ok = 0;
    }
    B32 collided = 0;
    if (ok && !full) {
      groups = 0;
      for (U32 i = 0; i < cblocks; i++) groups += bcounts[i];
      /* tuple keys read every group back to see that no two keys shared a
       * code; otherwise at most cap groups */
      U64 take = (T || groups < cap) ? groups : cap;
      free(recs);
      recs = (U64*)malloc((take ? take : 1) * 8 * rec_words);
      if (!recs) {
        ok = 0;
      } else if (take) {
        /* on a batch, so a failed read fails the query instead of leaving
         * zeros in the records */
        GPU_Batch* rb = gpu_batch_begin(0, take * 8 * rec_words);
        gpu_batch_buffer_read(rb, out, recs, take * 8 * rec_words);
        ok = gpu_batch_end(rb);
      }
      for (U64 g = 0; ok && T && g < take && !collided; g++) {
        const U64* r = recs + g * rec_words;
        for (U32 i = 0; i < T; i++)
          if (r[2 + 3 * E + i] != r[2 + 3 * E + T + i]) collided = 1;
      }
    }
    free(bcounts);
    qe_scratch_release(table);
    qe_scratch_release(bc);
    qe_scratch_release(out);
    if (!ok) break;
    if (!full && !collided) break;
    if (full) log2++;
    if (collided) seed++;
    if (attempt == 5 || log2 > 30) {
      qe_set_error(full ? "hash group by: table of 2^%u slots full" : "hash group by: key hash collisions at 2^%u slots",
                   log2);
      err_set = 1;
      ok = 0;
    }
  }
  qe_scratch_release(leaf_buf);
  qe_scratch_release(skip_buf);
  qe_scratch_release(flag_buf);
  qe_scratch_release(key_buf);
  if (!ok) {
    free(recs);
    if (!err_set) qe_set_error("aggregate_hash: %s", gpu_hip_last_error());
    return 0;
  }
  U64 take = groups < cap ? groups : cap;
  const B32 raw = qe_t_raw_int_sums;
  for (U64 g = 0; g < take; g++) {
    const U64* r = recs + g * rec_words;
    if (out_keys) {
      F64* kv = out_keys + g * K;
      if (mode == 0) {
        if (K) kv[0] = qe_hash_key_to_f64(key_type, r[0]);
      } else if (mode == 1) {
        qe_group_digits(keys, r[0], kv);
      } else {
        for (U32 i = 0; i < K; i++) kv[i] = qe_hash_key_to_f64(keys->type[i], r[2 + 3 * E + i]);
      }
    }
    U64 cnt = r[1];
    for (U32 e = 0; e < p->expr_count; e++) {
      S32 a = p->expr_arg[e];
      F64 sum = 0.0, mn = 0.0, mx = 0.0;
      if (a >= 0) {
        U32 ty = t->columns[p->arg_col[a]].type;
        U64 sv = r[2 + a];
        if (qe_type_int32(ty)) sum = (F64)(S64)sv;
        else memcpy(&sum, &sv, 8);
        mn = qe_f64_from_order_code(r[2 + E + a]);
        mx = qe_f64_from_order_code(r[2 + 2 * E + a]);
        if (qe_type_int32(ty)) {
          out_cells[g * p->expr_count + e] = qe_agg_cell_int(p->funcs[e], cnt, (S64)sv, mn, mx, raw);
          continue;
        }
      }
      out_cells[g * p->expr_count + e] = qe_agg_cell(p->funcs[e], cnt, sum, mn, mx);
    }
  }
  free(recs);
  *out_group_count = groups;
  if (trace) {
    trace->path = QE_AggPath_Hash;
    trace->kernel_count = 3;
    trace->rows_scanned = rows;
    trace->rows_after = groups;
    U64 bytes = 0;
    for (U32 sl = 0; sl < slots.count; sl++) bytes += rows * t->columns[slots.column[sl]].width;
    trace->bytes_read = bytes;
    snprintf(trace->kernel, sizeof(trace->kernel), "aggregate_hash");
  }
  return 1;
}

/* ---------------- entry point ---------------- */
static B32 qe_aggregate_run(QE_Table* table, const QE_IRNode* where, U32 where_count, const S32* group_cols,
                            U32 group_count, const U32* funcs, const S32* arg_cols, U32 expr_count, U64 cap,
                            U64* out_group_count, F64* out_keys, U64* out_cells, QE_Trace* trace)
{
  U64 t0, p0;
  qe_trace_begin(trace, &t0, &p0);
  if (!table || !funcs || !arg_cols || !out_group_count || !out_cells || (group_count && !group_cols)) {
    qe_set_error("qe_aggregate: null argument");
    return 0;
  }
  QE_GroupKeys keys;
  if (!qe_group_keys(table, group_cols, group_count, &keys)) return 0;
  const U32 K = keys.count;
  if (!qe_where_non_null(table, where, where_count, "qe_aggregate")) return 0;
  for (U32 i = 0; i < K; i++)
    if (!qe_require_non_null(table, (S32)keys.col[i], "qe_aggregate")) return 0;
  for (U32 e = 0; e < expr_count; e++)
    if (!qe_require_non_null(table, arg_cols[e], "qe_aggregate")) return 0;
  /* host-backed rows: streamed per chunk (QE_AGG_STREAM, query_exec.c:
   * 8842-9110, :10379), or with it 0 read in place over the host link */
  if (table->host_backed) {
    if (qe_settings_u64("QE_AGG_STREAM", 1) != 0)
      return qe_stream_aggregate(table, where, where_count, group_cols, group_count, funcs, arg_cols, expr_count, cap,
                                 out_group_count, out_keys, out_cells, trace);
    if (!qe_require_device(table, "qe_aggregate")) return 0;
  }
  QE_AggPlan plan;
  if (!qe_agg_plan(table, funcs, arg_cols, expr_count, &plan)) return 0;
  if (trace) trace->rows_before = table->row_count;

  /* WHERE: an AND / OR of column-literal comparisons, IN lists and
   * IS [NOT] NULL over at most 8 leaves fuses into the aggregate kernels */
  QE_Where wh;
  memset(&wh, 0, sizeof(wh));
  B32 general_where = 0;
  if (where && where_count) {
    static __thread QE_Bytecode prog;
    if (!qe_where_build(table, where, where_count, &wh)) {
      /* any other WHERE runs as a scan whose row ids stay on the device
       * and feed the hash aggregate */
      if (!qe_bytecode_program_build(table, where, where_count, &prog)) {
        qe_set_error("qe_aggregate: %s", prog.error);
        return 0;
      }
      general_where = 1;
    }
    /* QE_FUSE_AGGREGATE (query_exec.c:9892) / QE_FUSE_GROUP_BY (:9447) 0:
     * the WHERE runs as its own scan whose row ids feed the aggregate */
    if (!general_where && wh.leaf_count &&
        qe_settings_u64(K ? "QE_FUSE_GROUP_BY" : "QE_FUSE_AGGREGATE", 1) == 0)
      general_where = 1;
  }
  const QE_Where* wp = general_where ? 0 : &wh;
  if (general_where) {
    GPU_Buffer* ids = 0;
    U64 n = 0;
    if (!qe_scan_filter_device(table, where, where_count, &ids, &n)) return 0;
    QE_Prune all;
    memset(&all, 0, sizeof(all));
    all.row_end = table->row_count;
    U64 groups = 0;
    B32 ok = 1;
    if (n) ok = qe_aggregate_hash(table, &keys, 0, &all, ids, n, &plan, cap, &groups, out_keys, out_cells, trace);
    qe_scratch_release(ids);
    if (!ok) return 0;
    if (K == 0 && groups == 0) {
      /* an ungrouped aggregate over no rows is one row of zeros
       * (query_exec.c:10335-10346) */
      if (cap >= 1)
        for (U32 e = 0; e < expr_count; e++) out_cells[e] = qe_agg_cell(plan.funcs[e], 0, 0.0, 0.0, 0.0);
      groups = 1;
    }
    *out_group_count = groups;
    if (trace) {
      trace->path = QE_AggPath_Hash;
      trace->rows_after = groups;
    }
    qe_trace_end(trace, t0, p0);
    return 1;
  }
  QE_Leaf prune_leaves[32];
  U32 prune_count = qe_collect_prune_leaves(where, where_count, prune_leaves, 32);
  QE_Prune prune;
  qe_zonemap_prune(table, prune_leaves, prune_count, &prune);

  B32 ok = 1;
  if (K == 0) {
    if (cap < 1) {
      *out_group_count = 1;
      qe_prune_free(&prune);
      return 1;
    }
    if (prune.zones_left == 0 || table->row_count == 0) {
      for (U32 e = 0; e < expr_count; e++) out_cells[e] = qe_agg_cell(plan.funcs[e], 0, 0.0, 0.0, 0.0);
      if (trace) trace->path = prune.zones_left == 0 ? QE_AggPath_FusedGlobal : QE_AggPath_Global;
    } else {
      ok = qe_aggregate_global(table, wp, &prune, &plan, out_cells, trace);
    }
    if (ok) *out_group_count = 1;
    qe_prune_free(&prune);
    qe_trace_end(trace, t0, p0);
    return ok;
  }

  if (prune.zones_left == 0 || table->row_count == 0) {
    qe_prune_free(&prune);
    *out_group_count = 0;
    qe_trace_end(trace, t0, p0);
    return 1;
  }
  /* dense LDS tables when every key is a 32-bit integer whose range (from
   * the zone map, query_exec.c:7636-7647) is small, at most three keys
   * whose ranges multiply to a small range (digits, query_exec.c:9462-9500),
   * and the args are 32-bit integers; the hash table otherwise */
  const U64 span = qe_group_keys_span(&keys);
  U64 max_range = qe_settings_u64("QE_AGG_DENSE_MAX_RANGE", 65536);
  B32 dense = qe_settings_u64("QE_AGG_DENSE", 1) != 0 && K <= QE_DENSE_MAX_KEYS && span != 0 && span <= max_range;
  /* arguments wider than 32-bit integers take the wide tables
   * (QE_AGG_DENSE_WIDE, default 1), else the hash table */
  const B32 wide_args = qe_dense_wide_args(table, plan.arg_col, plan.arg_count);
  if (wide_args && qe_settings_u64("QE_AGG_DENSE_WIDE", 1) == 0) dense = 0;
  /* args that are key columns fold into the counts: SUM = key * count,
   * MIN = MAX = key (exact for whole-number keys) */
  S32 arg_key[QE_AGG_MAX_ARGS];
  B32 fold = qe_settings_u64("QE_AGG_KEY_FOLD", 1) != 0;
  for (U32 a = 0; a < plan.arg_count; a++) {
    arg_key[a] = -1;
    for (U32 i = 0; i < K; i++)
      if (plan.arg_col[a] == keys.col[i]) arg_key[a] = (S32)i;
    if (arg_key[a] < 0) fold = 0;
  }
  /* the block's LDS tables must fit: counts u32[R], sums u64[E][R],
   * min/max u32[E][R]; or the narrow tables of one argument */
  if (dense) {
    const U32 e = fold ? 0 : plan.arg_count;
    if (wide_args && !fold) {
      if (qe_dense_lds_wide((U32)span, e, plan.need_sum, plan.need_minmax) > gpu_hip_max_lds_bytes()) dense = 0;
    } else {
      const U64 lds = qe_dense_lds((U32)span, e, plan.need_sum, plan.need_minmax, 0);
      if (lds > gpu_hip_max_lds_bytes() &&
          (fold || !qe_dense_narrow_ok(table, plan.arg_col, plan.arg_count, plan.need_sum, plan.need_minmax, (U32)span)))
        dense = 0;
    }
  }
  if (!dense) {
    U64 groups = 0;
    ok = qe_aggregate_hash(table, &keys, wp, &prune, 0, 0, &plan, cap, &groups, out_keys, out_cells, trace);
    qe_prune_free(&prune);
    if (!ok) return 0;
    *out_group_count = groups;
    qe_trace_end(trace, t0, p0);
    return 1;
  }
  const U32 range = (U32)span;
  QE_DenseTotals tot;
  memset(&tot, 0, sizeof(tot));
  if (fold) ok = qe_dense_run(table, &keys, 0, 0, 0, 0, wp, &prune, &tot, trace);
  else
    ok = qe_dense_run(table, &keys, plan.arg_col, plan.arg_count, plan.need_sum, plan.need_minmax, wp, &prune, &tot,
                      trace);
  qe_prune_free(&prune);
  if (!ok) return 0;
  /* groups in key order (query_exec.c:8006-8039) */
  U64 g = 0;
  const B32 raw = qe_t_raw_int_sums;
  F64 kv[QE_AGG_MAX_GROUP_COLS];
  for (U32 i = 0; i < range; i++) {
    U64 cnt = tot.count[i];
    if (!cnt) continue;
    if (g < cap) {
      qe_group_digits(&keys, i, kv);
      if (out_keys)
        for (U32 k = 0; k < K; k++) out_keys[g * K + k] = kv[k];
      for (U32 e = 0; e < expr_count; e++) {
        S32 a = plan.expr_arg[e];
        F64 sum = 0.0, mn = 0.0, mx = 0.0;
        S64 isum = 0;
        if (a >= 0) {
          if (fold) {
            const F64 v = kv[arg_key[a]];
            isum = (S64)v * (S64)cnt;
            sum = v * (F64)cnt;
            mn = mx = v;
          } else {
            U32 ty = table->columns[plan.arg_col[a]].type;
            U64 s = tot.sum[(U64)a * range + i];
            if (tot.wide) {
              /* f64 sums for the wider types, min/max as f64 order keys */
              mn = qe_f64_from_order_code(tot.mn64[(U64)a * range + i]);
              mx = qe_f64_from_order_code(tot.mx64[(U64)a * range + i]);
              if (!qe_type_int32(ty)) {
                memcpy(&sum, &s, 8);
                out_cells[g * expr_count + e] = qe_agg_cell(plan.funcs[e], cnt, sum, mn, mx);
                continue;
              }
              isum = (S64)s;
            } else {
              isum = (S64)s;
              sum = ty == CQL_ColumnType_I32 ? (F64)(S64)s : (F64)s;
              U32 lo = tot.mn[(U64)a * range + i], hi = tot.mx[(U64)a * range + i];
              mn = ty == CQL_ColumnType_I32 ? (F64)(S32)(lo ^ 0x80000000u) : (F64)lo;
              mx = ty == CQL_ColumnType_I32 ? (F64)(S32)(hi ^ 0x80000000u) : (F64)hi;
            }
          }
          out_cells[g * expr_count + e] = qe_agg_cell_int(plan.funcs[e], cnt, isum, mn, mx, raw);
          continue;
        }
        out_cells[g * expr_count + e] = qe_agg_cell(plan.funcs[e], cnt, sum, mn, mx);
      }
    }
    g++;
  }
  qe_dense_totals_free(&tot);
  *out_group_count = g;
  if (trace) {
    trace->path = (wp && wp->leaf_count) ? QE_AggPath_FusedDense : (fold ? QE_AggPath_DenseKeyFold : QE_AggPath_Dense);
    trace->rows_after = g;
  }
  qe_trace_end(trace, t0, p0);
  return 1;
}

B32 qe_aggregate(QE_Table* table, const QE_IRNode* where, U32 where_count, const S32* group_cols, U32 group_count,
                 const U32* funcs, const S32* arg_cols, U32 expr_count, U64 cap, U64* out_group_count, F64* out_keys,
                 U64* out_cells, QE_Trace* trace)
{
  B32 ok = qe_aggregate_run(table, where, where_count, group_cols, group_count, funcs, arg_cols, expr_count, cap,
                            out_group_count, out_keys, out_cells, trace);
  /* chunk partials of a streamed aggregate carry raw integer sums; the
   * streamed operator checks its merged result instead */
  if (ok && !qe_t_raw_int_sums) {
    const U64 rows = *out_group_count < cap ? *out_group_count : cap;
    ok = qe_check_sum_exact(table, funcs, arg_cols, expr_count, out_cells, rows);
  }
  return ok;
}
