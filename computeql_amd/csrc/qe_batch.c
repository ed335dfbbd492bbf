/*
 * qe_batch.c - batched shared scan: many aggregate queries over one table
 * in one pass over the union of their columns.
 *
 * The reference runs one query per kernel (fused WHERE + aggregate,
 * query_exec.c:9442-10253 with query_kernel.c:1433-1600) and serialises
 * queries behind one mutex (application.h:34).  Here a batch of queries is
 * compiled into a program for the shared_scan kernel: each block walks row
 * tiles and evaluates every query of the launch on its rows, so a column is
 * read from HBM once per launch however many queries touch it.  Results
 * are decoded exactly as qe_aggregate decodes its own (qe_agg_cell), with
 * integer sums exact and f64 sums in a different (stated) order: for an
 * ungrouped query, a lane's rows of a tile, then the wave (DPP tree), then
 * the block's word across tiles in tile order; for a grouped query, the
 * rows' LDS atomic adds in arrival order; then per word the blocks, lane l
 * of the reducing wave folding blocks l, l + 64, ... in order and the wave
 * combining its lanes (DPP tree).
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdlib.h>
#include <sys/stat.h>

#include "qe_internal.h"

#define QE_BATCH_REC_WORDS 14
#define QE_BATCH_MAX_LAUNCH_QUERIES 256

typedef struct QE_BatchPlanned {
  U32 index; /* into the caller's array */
  QE_AggPlan plan;
  QE_Where where;
  S32 group_col;
  U32 key_min, range;
  U32 acc_off, acc_words;
  U32 reps, lds_off, lds_words; /* LDS replicas of the words (grouped queries) */
} QE_BatchPlanned;

/* LDS replicas of a grouped query's words: lanes of the owning wave whose
 * rows fall in the same group contend on one LDS word; with `reps` copies
 * (lane l adding into copy l mod reps) about QE_BATCH_REPLICA_LANES (32)
 * active lanes spread over range x reps words.  At most 16 copies and
 * 16 KB per query. */
static U32 qe_batch_reps(U32 range, U32 words)
{
  U64 lanes = qe_settings_u64("QE_BATCH_REPLICA_LANES", 32);
  U32 reps = 1;
  while (reps < 16 && (U64)range * reps < lanes) reps <<= 1;
  while (reps > 1 && (U64)words * reps * 8 > 16384) reps >>= 1;
  return reps;
}

static B32 qe_batch_int32(U32 t) { return t == CQL_ColumnType_U32 || t == CQL_ColumnType_I32; }

static F64 qe_batch_f64_from_code(U64 key)
{
  U64 b = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
  F64 d;
  memcpy(&d, &b, 8);
  return d;
}

/* can this query run in a shared scan?  fills `out` when it can */
static B32 qe_batch_plan_query(QE_Table* t, const QE_BatchQuery* q, QE_BatchPlanned* out)
{
  memset(out, 0, sizeof(*out));
  if (!q->funcs || !q->arg_cols || !q->out_group_count || !q->out_cells) return 0;
  if (!qe_agg_plan(t, q->funcs, q->arg_cols, q->expr_count, &out->plan)) return 0;
  if (out->plan.arg_count > 2) return 0;
  /* a nullable argument or key goes to qe_aggregate, which refuses it as
   * the reference does (query_exec.c:10312-10333); nullable columns in
   * the WHERE leaves run here, a NULL row failing its comparison */
  for (U32 a = 0; a < out->plan.arg_count; a++)
    if (t->columns[out->plan.arg_col[a]].has_nulls) return 0;
  /* several GROUP BY keys run on qe_aggregate */
  if (q->group_count > 1) return 0;
  const S32 group_col = q->group_count ? q->group_cols[0] : -1;
  if (group_col >= 0 && (U32)group_col < t->column_count && t->columns[group_col].has_nulls) return 0;
  /* any AND / OR / IN WHERE of at most 8 leaves and 16 terms */
  if (q->where && q->where_count && !qe_where_build(t, q->where, q->where_count, &out->where)) return 0;
  out->group_col = group_col;
  out->range = 1;
  if (group_col >= 0) {
    if ((U32)group_col >= t->column_count) return 0;
    QE_Column* k = &t->columns[group_col];
    if (!qe_batch_int32(k->type) || !k->zonemap_valid || t->row_count == 0) return 0;
    F64 span = k->max_value - k->min_value + 1.0;
    if (span < 1.0 || span > (F64)qe_settings_u64("QE_BATCH_MAX_GROUPS", 1024)) return 0;
    out->range = (U32)span;
    out->key_min = k->type == CQL_ColumnType_I32 ? ((U32)(S32)k->min_value ^ 0x80000000u) : (U32)k->min_value;
  }
  out->acc_words = out->range * (1 + 3 * out->plan.arg_count);
  out->reps = group_col >= 0 ? qe_batch_reps(out->range, out->acc_words) : 1;
  out->lds_words = out->acc_words * out->reps;
  return 1;
}

/* words of a query's record: the fixed part, 6 per leaf, the terms */
static U32 qe_batch_rec_words(const QE_BatchPlanned* q)
{
  return QE_BATCH_REC_WORDS + 6 * q->where.leaf_count + (q->where.term_count + 3) / 4;
}

/* a query's relative cost per tile in the shared scan: its leaves, and
 * per argument the LDS atomics (grouped) or wave reductions (ungrouped)
 * of its sum and min/max */
static U32 qe_batch_cost(const QE_BatchPlanned* q)
{
  const U32 per_arg = (q->plan.need_sum ? 1u : 0u) + (q->plan.need_minmax ? 2u : 0u);
  U32 c = 2 * q->where.leaf_count + q->where.term_count;
  if (q->group_col >= 0) c += 4 + 4 * q->plan.arg_count * per_arg;
  else c += 2 + 2 * q->plan.arg_count * per_arg;
  return c;
}

/* staged columns of a launch: the union of the columns its queries read,
 * in first-use order; returns the stage index of `column` (adding it).
 * The null bitmaps of the nullable columns its WHERE leaves read are staged
 * beside them (`nulls`, in first-use order). */
typedef struct QE_BatchStage {
  U32 count;
  U32 column[QE_MAX_COLUMNS];
  U32 nulls;
  U32 null_column[QE_MAX_COLUMNS];
} QE_BatchStage;

static U32 qe_batch_stage_of(QE_BatchStage* st, U32 column)
{
  for (U32 i = 0; i < st->count; i++)
    if (st->column[i] == column) return i;
  st->column[st->count] = column;
  return st->count++;
}

static U32 qe_batch_null_of(QE_BatchStage* st, U32 column)
{
  for (U32 i = 0; i < st->nulls; i++)
    if (st->null_column[i] == column) return i;
  st->null_column[st->nulls] = column;
  return st->nulls++;
}

static void qe_batch_stage_query(const QE_Table* t, QE_BatchStage* st, const QE_BatchPlanned* q)
{
  for (U32 l = 0; l < q->where.leaf_count; l++) {
    qe_batch_stage_of(st, q->where.leaf[l].column);
    if (t->columns[q->where.leaf[l].column].has_nulls) qe_batch_null_of(st, q->where.leaf[l].column);
  }
  if (q->group_col >= 0) qe_batch_stage_of(st, (U32)q->group_col);
  for (U32 a = 0; a < q->plan.arg_count; a++) qe_batch_stage_of(st, q->plan.arg_col[a]);
}

/* the staged columns' bytes in a stage buffer for tiles of 256 * items rows
 * (each a multiple of 1 KB) */
static U64 qe_batch_stage_column_bytes(QE_Table* t, const QE_BatchStage* st, U32 items)
{
  U64 b = 0;
  for (U32 i = 0; i < st->count; i++) b += (U64)QE_WG * items * t->columns[st->column[i]].width;
  return b;
}

/* 1 KB chunks of staged null bitmaps: a bitmap is tile / 8 bytes, a chunk
 * holds 8192 / tile of them */
static U32 qe_batch_null_chunks(const QE_BatchStage* st, U32 items)
{
  const U32 per = 8192u / (QE_WG * items);
  return (st->nulls + per - 1) / per;
}

/* bytes of one stage buffer for tiles of 256 * items rows: the columns,
 * then the null bitmaps */
static U64 qe_batch_stage_bytes(QE_Table* t, const QE_BatchStage* st, U32 items)
{
  return qe_batch_stage_column_bytes(t, st, items) + 1024ull * qe_batch_null_chunks(st, items);
}

/* the dynamic LDS a shared-scan block may take: every instance keeps the
 * same static words (its query counters) beside the stage and tables */
static U32 qe_batch_max_lds(void) { return gpu_hip_kernel_max_lds_bytes(qe_kernel("shared_scan_256_1")); }

static U64 qe_batch_lds_n(U64 W, U64 stage_bytes, U32 buffers) { return ((W * 8 + 15) & ~15ull) + buffers * stage_bytes; }
static U64 qe_batch_lds(U64 W, U64 stage_bytes) { return qe_batch_lds_n(W, stage_bytes, 1); }

/* program words of one launch (layout: qk_library.hip, shared scan) */
static U32* qe_batch_program(QE_Table* t, QE_BatchPlanned* qs, U32 n, U32 items, U32* out_words, U32* out_w,
                             U32* out_wl, U64* out_stage_bytes)
{
  if (n > QE_BATCH_MAX_LAUNCH_QUERIES) return 0; /* launches are split below that */
  QE_BatchStage st;
  memset(&st, 0, sizeof(st));
  U32 W = 0, WL = 0;
  for (U32 i = 0; i < n; i++) {
    qs[i].acc_off = W;
    W += qs[i].acc_words;
    qs[i].lds_off = WL;
    WL += qs[i].lds_words;
    qe_batch_stage_query(t, &st, &qs[i]);
  }
  U32 rec_total = 0;
  for (U32 i = 0; i < n; i++) rec_total += qe_batch_rec_words(&qs[i]);
  const U32 hdr = 12;
  /* the chunk table: a tile's stage as 1 KB chunks, the columns' then the
   * null bitmaps' */
  U32 chunks = 0;
  for (U32 u = 0; u < st.count; u++) chunks += (U32)((U64)QE_WG * items * t->columns[st.column[u]].width / 1024);
  const U32 null_chunks = qe_batch_null_chunks(&st, items);
  chunks += null_chunks;
  U32 words = hdr + n + 3 * st.count + rec_total + W + chunks + 2 * st.nulls;
  U32* p = (U32*)calloc(words, 4);
  if (!p) return 0;
  /* grouped queries walk only their passing rows (QE_BATCH_COMPACT, 1) */
  const B32 compact = qe_settings_u64("QE_BATCH_COMPACT", 1) != 0;
  U64 stage_bytes = qe_batch_stage_bytes(t, &st, items);
  p[0] = n;
  p[1] = W;
  p[2] = hdr + n + 3 * st.count + rec_total;
  p[3] = st.count;
  p[4] = (U32)stage_bytes;
  p[5] = WL;
  p[6] = p[2] + W;
  p[7] = chunks;
  /* the staged null bitmaps: count, their table (table column, byte
   * offset of its bitmap / 16), the bitmap area's byte offset in a stage
   * buffer and its chunks */
  const U32 bitmap_area = (U32)qe_batch_stage_column_bytes(t, &st, items);
  p[8] = st.nulls;
  p[9] = p[6] + chunks;
  p[10] = bitmap_area;
  p[11] = null_chunks;
  for (U32 i = 0; i < st.nulls; i++) {
    p[p[9] + 2 * i] = st.null_column[i];
    p[p[9] + 2 * i + 1] = (U32)(t->columns[st.null_column[i]].null_offset / 16);
  }
  {
    U32 f = 0;
    for (U32 u = 0; u < st.count; u++) {
      const U32 cu = (U32)((U64)QE_WG * items * t->columns[st.column[u]].width / 1024);
      for (U32 k = 0; k < cu; k++) p[p[6] + f++] = u | (k << 16);
    }
    for (U32 j = 0; j < null_chunks; j++) p[p[6] + f++] = 0x8000u | (j << 16);
  }
  U32 off = 0;
  for (U32 u = 0; u < st.count; u++) {
    U32 w = t->columns[st.column[u]].width;
    p[hdr + n + 3 * u] = st.column[u];
    p[hdr + n + 3 * u + 1] = w;
    p[hdr + n + 3 * u + 2] = off;
    off += QE_WG * items * w;
  }
  /* the records are listed most expensive first: the kernel's waves take
   * them in that order, so the long ones start first and the short ones
   * fill in (longest-first list scheduling) */
  U32 order[QE_BATCH_MAX_LAUNCH_QUERIES], cost[QE_BATCH_MAX_LAUNCH_QUERIES];
  for (U32 i = 0; i < n; i++) {
    order[i] = i;
    cost[i] = qe_batch_cost(&qs[i]);
  }
  for (U32 i = 1; i < n; i++) {
    const U32 o = order[i];
    U32 j = i;
    while (j > 0 && cost[order[j - 1]] < cost[o]) {
      order[j] = order[j - 1];
      j--;
    }
    order[j] = o;
  }
  U32 rank_of[QE_BATCH_MAX_LAUNCH_QUERIES];
  for (U32 r = 0; r < n; r++) rank_of[order[r]] = r;
  U32 at = hdr + n + 3 * st.count;
  for (U32 i = 0; i < n; i++) {
    QE_BatchPlanned* q = &qs[i];
    p[hdr + rank_of[i]] = at;
    U32* r = p + at;
    r[0] = q->where.leaf_count | (q->where.term_count << 8);
    r[1] = (q->group_col >= 0 ? 1u : 0u) | (q->plan.need_sum ? 2u : 0u) | (q->plan.need_minmax ? 4u : 0u) |
           (q->group_col >= 0 && compact ? 8u : 0u);
    /* the key's and arguments' rows as byte offsets in a stage buffer */
    r[2] = q->group_col >= 0 ? p[hdr + n + 3 * qe_batch_stage_of(&st, (U32)q->group_col) + 2] : 0;
    r[3] = q->group_col >= 0 ? t->columns[q->group_col].type : 0;
    r[4] = q->key_min;
    r[5] = q->range;
    r[6] = q->plan.arg_count;
    for (U32 a = 0; a < q->plan.arg_count; a++) {
      r[7 + 2 * a] = p[hdr + n + 3 * qe_batch_stage_of(&st, q->plan.arg_col[a]) + 2];
      r[8 + 2 * a] = t->columns[q->plan.arg_col[a]].type;
    }
    r[11] = q->acc_off;
    r[12] = q->lds_off;
    r[13] = q->reps;
    if (q->where.leaf_count) {
      /* the staged columns hold full-width values: no encodings */
      U32 slot[QE_SCAN_MAX_LEAVES], words[QE_WHERE_WORDS], packed = 0;
      for (U32 l = 0; l < q->where.leaf_count; l++) slot[l] = qe_batch_stage_of(&st, q->where.leaf[l].column);
      qe_where_words(t, &q->where, slot, 1, 0, words, &packed);
      const U32 L = q->where.leaf_count;
      /* word 0 of a leaf: its column's byte offset in a stage buffer (in
       * place of the stage index); word 5 of a nullable column's leaf: 1 +
       * its staged bitmap's byte offset in a stage buffer / 16 */
      for (U32 l = 0; l < L; l++) words[6 * l] = p[hdr + n + 3 * slot[l] + 2];
      for (U32 l = 0; l < L; l++)
        if (words[6 * l + 5])
          words[6 * l + 5] = 1u + (bitmap_area + qe_batch_null_of(&st, q->where.leaf[l].column) * (QE_WG * items / 8)) / 16;
      memcpy(&r[QE_BATCH_REC_WORDS], words, 24 * L);
      for (U32 k = 0; k < (q->where.term_count + 3) / 4; k++) r[QE_BATCH_REC_WORDS + 6 * L + k] = words[6 * QE_SCAN_MAX_LEAVES + k];
    }
    at += qe_batch_rec_words(q);
    /* op of every accumulator word: count add; per arg sum (integer or
     * f64 add), min, max */
    U32* ops = p + p[2] + q->acc_off;
    U32 R = q->range;
    for (U32 g = 0; g < R; g++) ops[g] = 0;
    for (U32 a = 0; a < q->plan.arg_count; a++) {
      U32 sum_op = qe_batch_int32(t->columns[q->plan.arg_col[a]].type) ? 0u : 1u;
      for (U32 g = 0; g < R; g++) {
        ops[R + 3 * a * R + g] = sum_op;
        ops[R + 3 * a * R + R + g] = 2;
        ops[R + 3 * a * R + 2 * R + g] = 3;
      }
    }
  }
  *out_words = words;
  *out_w = W;
  *out_wl = WL;
  *out_stage_bytes = stage_bytes;
  return p;
}

/* the query's result from its accumulator words */
static void qe_batch_decode(QE_Table* t, const QE_BatchPlanned* q, const QE_BatchQuery* bq, const U64* acc)
{
  const QE_AggPlan* p = &q->plan;
  const U64* a = acc + q->acc_off;
  U32 R = q->range;
  U64 g = 0;
  for (U32 i = 0; i < R; i++) {
    U64 cnt = a[i];
    if (q->group_col >= 0 && !cnt) continue;
    if (g < bq->cap) {
      if (bq->out_keys && q->group_col >= 0) {
        U32 code = q->key_min + i;
        bq->out_keys[g] = t->columns[q->group_col].type == CQL_ColumnType_I32 ? (F64)(S32)(code ^ 0x80000000u) : (F64)code;
      }
      for (U32 e = 0; e < p->expr_count; e++) {
        S32 ai = p->expr_arg[e];
        F64 sum = 0.0, mn = 0.0, mx = 0.0;
        if (ai >= 0 && cnt > 0) {
          const U64* s = a + R + 3 * (U64)ai * R;
          if (qe_batch_int32(t->columns[p->arg_col[ai]].type)) sum = (F64)(S64)s[i];
          else memcpy(&sum, &s[i], 8);
          mn = qe_batch_f64_from_code(s[R + i]);
          mx = qe_batch_f64_from_code(s[2 * R + i]);
        }
        bq->out_cells[g * p->expr_count + e] = qe_agg_cell(p->funcs[e], cnt, sum, mn, mx);
      }
    }
    g++;
  }
  *bq->out_group_count = g;
}

/* shared_scan instances: block size and rows per thread staged, in order
 * of preference (more rows per lane per query evaluation first) */
static const struct {
  U32 wg, items;
} qe_batch_shapes[] = {{1024, 1}, {512, 2}, {512, 1}, {256, 4}, {256, 2}, {256, 1}};
#define QE_BATCH_SHAPES (sizeof(qe_batch_shapes) / sizeof(qe_batch_shapes[0]))

/* one shared-scan launch over queries qs[0..n) */
static B32 qe_batch_launch(QE_Table* t, QE_BatchPlanned* qs, U32 n, QE_BatchQuery* user, const QE_BatchConfig* cfg,
                           QE_BatchTrace* trace)
{
  const U32 max_items = cfg->items, bpc = cfg->blocks_per_cu;
  /* the first instance (block size, rows per thread) whose stage and
   * accumulators fit one block's LDS; QE_BATCH_WG / QE_BATCH_ITEMS pin
   * the shape, QE_BATCH_DOUBLE_BUFFER 1 asks for two stage buffers (the
   * default 2 takes them when they fit) */
  QE_BatchStage st;
  memset(&st, 0, sizeof(st));
  U64 Wq = 0;
  for (U32 i = 0; i < n; i++) {
    qe_batch_stage_query(t, &st, &qs[i]);
    Wq += qs[i].lds_words;
  }
  U64 want_wg = cfg->wg;
  U64 dbl_mode = cfg->buffers;
  U32 max_lds = qe_batch_max_lds();
  U32 wg = 256, items = 1, buffers = 1;
  B32 found = 0;
  for (U32 i = 0; i < QE_BATCH_SHAPES && !found; i++) {
    if (want_wg && qe_batch_shapes[i].wg != want_wg) continue;
    if (qe_batch_shapes[i].items > max_items) continue;
    U64 sb = qe_batch_stage_bytes(t, &st, qe_batch_shapes[i].items * qe_batch_shapes[i].wg / QE_WG);
    U32 nb = dbl_mode == 1 ? 2 : 1;
    if (qe_batch_lds_n(Wq, sb, nb) > max_lds) continue;
    wg = qe_batch_shapes[i].wg;
    items = qe_batch_shapes[i].items;
    buffers = (dbl_mode == 1 || (dbl_mode == 2 && qe_batch_lds_n(Wq, sb, 2) <= max_lds)) ? 2 : 1;
    found = 1;
  }
  U32 words = 0, W = 0, WL = 0;
  U64 stage_bytes = 0;
  U32* prog = qe_batch_program(t, qs, n, items * wg / QE_WG, &words, &W, &WL, &stage_bytes);
  if (!prog) {
    qe_set_error("qe_aggregate_batch: out of memory");
    return 0;
  }
  /* stage mode (push slot 6): 1 two buffers copied into LDS, 0 one, 2 one
   * filled from registers that take the next tile's loads while a tile is
   * evaluated (1024-thread blocks, a stage of at most 160 chunks of 1 KB;
   * QE_BATCH_STAGE_REGS, default 1) */
  U32 stage_mode = buffers == 2 ? 1u : 0u;
  if (wg == 1024 && stage_bytes <= 160u * 1024u && (stage_bytes & 1023u) == 0 &&
      qe_settings_u64("QE_BATCH_STAGE_REGS", 1) && dbl_mode != 1) {
    stage_mode = 2;
    buffers = 1;
  }
  U64 lds = qe_batch_lds_n(WL, stage_bytes, buffers);
  U64 rows = t->row_count;
  U64 tile = (U64)wg * items;
  U64 tiles = (rows + tile - 1) / tile;
  U64 cus = gpu_hip_compute_units();
  U64 fit = lds ? (U64)max_lds / lds : bpc;
  U64 wave_fit = 32 / (wg / 64); /* 8 waves per SIMD */
  if (fit > wave_fit) fit = wave_fit;
  if (fit < 1) fit = 1;
  U64 blocks = cus * (bpc < fit ? bpc : fit);
  if (blocks > tiles) blocks = tiles ? tiles : 1;
  /* the register-staged form has an instance of its own (no LDS-direct
   * copies in it, so no waits for them at its LDS atomics) */
  char kname[32];
  snprintf(kname, sizeof(kname), "shared_scan_%u_%u%s", wg, items, stage_mode == 2 ? "_r" : "");
  GPU_Kernel* k = qe_kernel(kname);
  GPU_Kernel* kr = qe_kernel("shared_reduce");
  GPU_Buffer* pb = qe_scratch("batch_program", (U64)words * 4);
  GPU_Buffer* cb = qe_scratch("batch_columns", (U64)QE_MAX_COLUMNS * 8);
  GPU_Buffer* part = qe_scratch("batch_partials", blocks * W * 8);
  GPU_Buffer* res = qe_scratch("batch_result", (U64)W * 8);
  U64* acc = (U64*)malloc((U64)W * 8 + 8);
  B32 ok = k && kr && pb && cb && part && res && acc;
  if (ok) {
    U64 colptr[QE_MAX_COLUMNS];
    memset(colptr, 0, sizeof(colptr));
    for (U32 c = 0; c < t->column_count; c++) colptr[c] = gpu_hip_buffer_device_address(t->columns[c].buffer);
    gpu_kernel_set_arg_buffer(k, 0, pb);
    gpu_kernel_set_arg_buffer(k, 1, part);
    gpu_kernel_set_arg_buffer(k, 2, cb);
    gpu_kernel_set_arg_u64(k, 0, rows);
    gpu_kernel_set_arg_u64(k, 5, 0);
    gpu_kernel_set_arg_u64(k, 6, stage_mode); /* push slot 6: how the stage is filled */
    gpu_hip_kernel_set_lds_bytes(k, (U32)lds);
    gpu_kernel_set_arg_buffer(kr, 0, pb);
    gpu_kernel_set_arg_buffer(kr, 1, part);
    gpu_kernel_set_arg_buffer(kr, 3, res);
    gpu_kernel_set_arg_u64(kr, 0, blocks);
    U32 rblocks = (W + QE_WG / 64 - 1) / (QE_WG / 64); /* one wave per word */
    /* QE_BATCH_GRAPH (default 1): the launch runs as a HIP graph that the
     * next batch of the same shape over the same table replays (only the
     * program bytes change); the shape includes the dynamic LDS size,
     * which the captured launch keeps */
    GPU_Batch* b;
    if (qe_settings_u64("QE_BATCH_GRAPH", 1)) {
      char key[160];
      snprintf(key, sizeof(key), "batch:%s:%p:%u:%u:%u:%llu:%llu", kname, (void*)t, words, W, stage_mode,
               (unsigned long long)blocks, (unsigned long long)lds);
      b = gpu_hip_batch_begin_graph(key);
    } else {
      b = gpu_batch_begin(0, (U64)W * 8);
    }
    gpu_batch_buffer_write(b, pb, prog, (U64)words * 4);
    gpu_batch_buffer_write(b, cb, colptr, (U64)t->column_count * 8);
    gpu_batch_kernel_execute(b, k, (U32)(blocks * wg), wg);
    gpu_batch_kernel_execute(b, kr, (rblocks ? rblocks : 1) * QE_WG, QE_WG);
    gpu_batch_buffer_read(b, res, acc, (U64)W * 8);
    ok = gpu_batch_end(b);
    gpu_hip_kernel_set_lds_bytes(k, 0);
    if (!ok) qe_set_error("qe_aggregate_batch: %s", gpu_hip_last_error());
  } else if (!k || !kr) {
    /* qe_kernel set the error */
  } else {
    qe_set_error("qe_aggregate_batch: out of memory");
  }
  if (ok) {
    for (U32 i = 0; i < n; i++) {
      QE_BatchQuery* bq = &user[qs[i].index];
      qe_batch_decode(t, &qs[i], bq, acc);
      /* a SUM reaching 2^53 fails this query (query_exec.c:7028-7049) */
      const U64 g = *bq->out_group_count;
      bq->ok = qe_check_sum_exact(t, bq->funcs, bq->arg_cols, bq->expr_count, bq->out_cells, g < bq->cap ? g : bq->cap);
      bq->fused = 1;
    }
    if (trace) {
      U8 used[QE_MAX_COLUMNS];
      memset(used, 0, sizeof(used));
      for (U32 i = 0; i < n; i++) {
        for (U32 l = 0; l < qs[i].where.leaf_count; l++) used[qs[i].where.leaf[l].column] = 1;
        for (U32 a = 0; a < qs[i].plan.arg_count; a++) used[qs[i].plan.arg_col[a]] = 1;
        if (qs[i].group_col >= 0) used[qs[i].group_col] = 1;
      }
      for (U32 c = 0; c < t->column_count; c++)
        if (used[c]) trace->bytes_read += rows * t->columns[c].width;
      trace->rows_scanned += rows;
      trace->launches++;
      trace->items = items;
      trace->blocks_per_cu = (U32)(blocks / (cus ? cus : 1));
    }
  }
  free(acc);
  free(prog);
  qe_scratch_release(pb);
  qe_scratch_release(cb);
  qe_scratch_release(part);
  qe_scratch_release(res);
  return ok;
}

/* the configuration the settings give (QE_BATCH_WG, QE_BATCH_ITEMS,
 * QE_BATCH_DOUBLE_BUFFER, QE_BATCH_BLOCKS_PER_CU, QE_BATCH_LDS_BYTES) */
static void qe_batch_config_from_settings(QE_BatchConfig* c)
{
  c->wg = (U32)qe_settings_u64("QE_BATCH_WG", 0);
  c->items = (U32)qe_settings_u64("QE_BATCH_ITEMS", 8);
  c->buffers = (U32)qe_settings_u64("QE_BATCH_DOUBLE_BUFFER", 2);
  c->blocks_per_cu = (U32)qe_settings_u64("QE_BATCH_BLOCKS_PER_CU", 2);
  c->lds_budget = qe_settings_u64("QE_BATCH_LDS_BYTES", 0);
}

/* the same query: WHERE nodes, keys, functions, arguments, capacity and
 * whether it asks for its keys */
static B32 qe_batch_same(const QE_BatchQuery* a, const QE_BatchQuery* b)
{
  if (a->where_count != b->where_count || a->group_count != b->group_count || a->expr_count != b->expr_count ||
      a->cap != b->cap || (a->out_keys != 0) != (b->out_keys != 0) || a->group_count > QE_AGG_MAX_GROUP_COLS)
    return 0;
  if (memcmp(a->group_cols, b->group_cols, a->group_count * sizeof(S32))) return 0;
  if (a->where_count && (!a->where || !b->where || memcmp(a->where, b->where, a->where_count * sizeof(QE_IRNode))))
    return 0;
  return memcmp(a->funcs, b->funcs, a->expr_count * 4) == 0 && memcmp(a->arg_cols, b->arg_cols, a->expr_count * 4) == 0;
}

/* FNV-1a over what qe_batch_same compares */
static U64 qe_batch_query_hash(const QE_BatchQuery* q)
{
  U64 h = 0xcbf29ce484222325ull;
#define QE_HASH_BYTES(p, n)                                                                                            \
  do {                                                                                                                 \
    const U8* b_ = (const U8*)(p);                                                                                     \
    for (U64 k_ = 0; k_ < (U64)(n); k_++) h = (h ^ b_[k_]) * 0x100000001b3ull;                                         \
  } while (0)
  QE_HASH_BYTES(&q->where_count, sizeof(q->where_count));
  const U32 want_keys = q->out_keys != 0;
  QE_HASH_BYTES(&want_keys, sizeof(want_keys));
  QE_HASH_BYTES(&q->group_count, sizeof(q->group_count));
  if (q->group_count <= QE_AGG_MAX_GROUP_COLS) QE_HASH_BYTES(q->group_cols, q->group_count * sizeof(S32));
  QE_HASH_BYTES(&q->expr_count, sizeof(q->expr_count));
  QE_HASH_BYTES(&q->cap, sizeof(q->cap));
  if (q->where_count && q->where) QE_HASH_BYTES(q->where, q->where_count * sizeof(QE_IRNode));
  if (q->expr_count) {
    QE_HASH_BYTES(q->funcs, q->expr_count * 4);
    QE_HASH_BYTES(q->arg_cols, q->expr_count * 4);
  }
#undef QE_HASH_BYTES
  return h;
}

/* rep_of[i] = the first earlier query identical to query i, or ~0u; an
 * open-addressing table of first occurrences keyed by the query hash keeps
 * this linear in the batch size */
static B32 qe_batch_find_repeats(const QE_BatchQuery* queries, U32 count, U32* rep_of)
{
  for (U32 i = 0; i < count; i++) rep_of[i] = ~0u;
  if (count < 2) return 1;
  U32 cap = 4;
  while (cap < 2 * count) cap <<= 1;
  U32* slot = (U32*)malloc(sizeof(U32) * cap);
  U64* hash = (U64*)malloc(sizeof(U64) * count);
  if (!slot || !hash) {
    free(slot);
    free(hash);
    return 0;
  }
  for (U32 i = 0; i < cap; i++) slot[i] = ~0u;
  for (U32 i = 0; i < count; i++) {
    hash[i] = qe_batch_query_hash(&queries[i]);
    for (U32 s = (U32)(hash[i] ^ (hash[i] >> 32)) & (cap - 1);; s = (s + 1) & (cap - 1)) {
      const U32 j = slot[s];
      if (j == ~0u) {
        slot[s] = i;
        break;
      }
      if (hash[j] == hash[i] && qe_batch_same(&queries[i], &queries[j])) {
        rep_of[i] = j;
        break;
      }
    }
  }
  free(slot);
  free(hash);
  return 1;
}

static void qe_batch_copy_result(const QE_BatchQuery* from, QE_BatchQuery* to)
{
  to->ok = from->ok;
  to->fused = from->fused;
  if (!from->ok) return;
  U64 g = *from->out_group_count;
  *to->out_group_count = g;
  U64 n = g < to->cap ? g : to->cap;
  if (to->out_keys && from->out_keys) memcpy(to->out_keys, from->out_keys, n * to->group_count * sizeof(F64));
  memcpy(to->out_cells, from->out_cells, n * to->expr_count * sizeof(U64));
}

/* A launch's queries as one shared scan, or one plain operator each: the
 * shared scan reads the union of their columns once but costs more per
 * query and row than a single-query kernel, which streams at the HBM rate.
 * Estimates: shared = max(union bytes / BW, max(queries, Q0) x rows / R)
 * and plain = the queries' bytes / BW + queries x launch, with BW
 * QE_BATCH_COST_BW_GBS (5500), R QE_BATCH_COST_SHARED_MROWS (600000 M
 * query-rows/s, measured on the 64-query mix) and launch
 * QE_BATCH_COST_LAUNCH_US (60).  Q0 (QE_BATCH_COST_MIN_QUERIES, 8): the
 * shared scan's waves take a tile's queries one per wave, so a launch of
 * few queries leaves most of a block's 16 waves idle; measured, one query
 * alone costs about what eight cost (a 3-leaf AND/OR count over 1B rows
 * 14.0 ms shared against 1.8 ms plain, GROUP BY region over 100M rows 1.5
 * against 0.17 ms). */
static B32 qe_batch_plain_cheaper(QE_Table* t, const QE_BatchPlanned* qs, U32 n)
{
  if (qe_settings_u64("QE_BATCH_COST", 1) == 0) return 0;
  const F64 bw = (F64)qe_settings_u64("QE_BATCH_COST_BW_GBS", 5500) * 1e9;
  const F64 rate = (F64)qe_settings_u64("QE_BATCH_COST_SHARED_MROWS", 600000) * 1e6;
  const F64 launch = (F64)qe_settings_u64("QE_BATCH_COST_LAUNCH_US", 60) * 1e-6;
  const F64 q0 = (F64)qe_settings_u64("QE_BATCH_COST_MIN_QUERIES", 8);
  const F64 rows = (F64)t->row_count;
  QE_BatchStage all;
  memset(&all, 0, sizeof(all));
  F64 plain = 0.0;
  for (U32 i = 0; i < n; i++) {
    QE_BatchStage one;
    memset(&one, 0, sizeof(one));
    qe_batch_stage_query(t, &one, &qs[i]);
    qe_batch_stage_query(t, &all, &qs[i]);
    F64 bytes = 0.0;
    for (U32 c = 0; c < one.count; c++) bytes += rows * t->columns[one.column[c]].width;
    plain += bytes / bw + launch;
  }
  F64 ubytes = 0.0;
  for (U32 c = 0; c < all.count; c++) ubytes += rows * t->columns[all.column[c]].width;
  F64 shared = ubytes / bw;
  const F64 qn = (F64)n > q0 ? (F64)n : q0;
  if (qn * rows / rate > shared) shared = qn * rows / rate;
  return plain < shared;
}

static B32 qe_aggregate_batch_run(QE_Table* table, QE_BatchQuery* queries, U32 count, const QE_BatchConfig* cfg_in,
                                  QE_BatchTrace* trace)
{
  U64 t0, p0;
  qe_trace_begin(0, &t0, &p0);
  if (!table || (!queries && count)) {
    qe_set_error("qe_aggregate_batch: null argument");
    return 0;
  }
  if (trace) memset(trace, 0, sizeof(*trace));
  QE_BatchConfig cfg = *cfg_in;
  cfg.items = cfg.items >= 8 ? 8 : (cfg.items >= 4 ? 4 : (cfg.items >= 2 ? 2 : 1));
  if (cfg.blocks_per_cu < 1) cfg.blocks_per_cu = 1;
  if (cfg.blocks_per_cu > 16) cfg.blocks_per_cu = 16;
  /* launches are packed so one block's LDS holds a 1024-row stage of
   * their columns and their accumulators (lds_budget, at most the LDS of
   * a CU; 0 = all of it) */
  U32 max_lds = qe_batch_max_lds();
  U64 lds_budget = cfg.lds_budget && cfg.lds_budget < max_lds ? cfg.lds_budget : max_lds;
  QE_BatchPlanned* planned = (QE_BatchPlanned*)calloc(count ? count : 1, sizeof(QE_BatchPlanned));
  if (!planned) {
    qe_set_error("qe_aggregate_batch: out of memory");
    return 0;
  }
  U32 np = 0, fallback = 0;
  B32 all_ok = 1;
  /* identical queries of the batch run once; the others copy the result
   * (QE_BATCH_DEDUPE, 1) */
  U32* rep_of = (U32*)malloc(sizeof(U32) * (count ? count : 1));
  if (!rep_of) {
    free(planned);
    qe_set_error("qe_aggregate_batch: out of memory");
    return 0;
  }
  for (U32 i = 0; i < count; i++) rep_of[i] = ~0u;
  if (qe_settings_u64("QE_BATCH_DEDUPE", 1) != 0 && !qe_batch_find_repeats(queries, count, rep_of)) {
    free(rep_of);
    free(planned);
    qe_set_error("qe_aggregate_batch: out of memory");
    return 0;
  }
  for (U32 i = 0; i < count; i++) {
    queries[i].ok = 0;
    queries[i].fused = 0;
    if (rep_of[i] != ~0u) continue;
    QE_BatchStage one;
    memset(&one, 0, sizeof(one));
    /* a host-backed table's queries each stream through the resident chunk
     * cache (qe_aggregate) rather than read its rows over the host link */
    if (!table->host_backed && qe_batch_plan_query(table, &queries[i], &planned[np]) &&
        (qe_batch_stage_query(table, &one, &planned[np]),
         qe_batch_lds(planned[np].lds_words, qe_batch_stage_bytes(table, &one, 1)) <= max_lds)) {
      planned[np].index = i;
      np++;
    } else {
      /* not shaped for the shared scan: the plain operator, same call */
      QE_BatchQuery* q = &queries[i];
      q->ok = qe_aggregate(table, q->where, q->where_count, q->group_cols, q->group_count, q->funcs, q->arg_cols,
                           q->expr_count, q->cap,
                           q->out_group_count, q->out_keys, q->out_cells, 0);
      if (!q->ok) all_ok = 0;
      fallback++;
    }
  }
  /* launches by column affinity: a launch takes, one at a time, the
   * remaining query that adds the fewest bytes to its stage (ties to the
   * earliest), while its accumulators and the stage buffers of
   * one-row-per-thread tiles fit the LDS budget; queries sharing columns
   * thus share launches and each launch reads a small union */
  {
    QE_BatchPlanned* ordered = (QE_BatchPlanned*)calloc(np ? np : 1, sizeof(QE_BatchPlanned));
    U8* taken = (U8*)calloc(np ? np : 1, 1);
    U32 done = 0;
    if (!ordered || !taken) {
      free(ordered);
      free(taken);
      free(planned);
      free(rep_of);
      qe_set_error("qe_aggregate_batch: out of memory");
      return 0;
    }
    while (done < np) {
      QE_BatchStage st;
      memset(&st, 0, sizeof(st));
      U64 words = 0;
      U32 in_launch = 0;
      for (;;) {
        U32 best = np;
        U64 best_add = ~0ull;
        U64 base = qe_batch_stage_bytes(table, &st, 4);
        for (U32 i = 0; i < np; i++) {
          if (taken[i]) continue;
          QE_BatchStage grown = st;
          qe_batch_stage_query(table, &grown, &planned[i]);
          U64 add = qe_batch_stage_bytes(table, &grown, 4) - base;
          if (in_launch && qe_batch_lds(words + planned[i].lds_words, base + add) > lds_budget) continue;
          if (add < best_add) {
            best_add = add;
            best = i;
          }
        }
        if (best == np || in_launch == QE_BATCH_MAX_LAUNCH_QUERIES) break;
        qe_batch_stage_query(table, &st, &planned[best]);
        words += planned[best].lds_words;
        taken[best] = 1;
        ordered[done++] = planned[best];
        in_launch++;
      }
      /* mark the launch boundary in the sequence */
      ordered[done - 1].acc_off = ~0u;
    }
    memcpy(planned, ordered, sizeof(QE_BatchPlanned) * np);
    free(ordered);
    free(taken);
  }
  U32 start = 0;
  while (start < np) {
    U32 end = start;
    while (planned[end].acc_off != ~0u) end++;
    end++;
    if (table->row_count == 0) {
      /* no rows: no groups, or the one all-zero row of an ungrouped
       * aggregate (query_exec.c:10335-10346) */
      for (U32 i = start; i < end; i++) {
        QE_BatchQuery* bq = &queries[planned[i].index];
        if (planned[i].group_col >= 0) {
          *bq->out_group_count = 0;
        } else {
          U64 acc[1 + 3 * 2];
          memset(acc, 0, sizeof(acc));
          planned[i].acc_off = 0;
          qe_batch_decode(table, &planned[i], bq, acc);
        }
        bq->ok = 1;
        bq->fused = 1;
      }
    } else if (qe_batch_plain_cheaper(table, planned + start, end - start)) {
      for (U32 i = start; i < end; i++) {
        QE_BatchQuery* q = &queries[planned[i].index];
        q->ok = qe_aggregate(table, q->where, q->where_count, q->group_cols, q->group_count, q->funcs, q->arg_cols,
                           q->expr_count, q->cap,
                             q->out_group_count, q->out_keys, q->out_cells, 0);
        q->fused = 0;
        if (!q->ok) all_ok = 0;
        fallback++;
      }
    } else if (!qe_batch_launch(table, planned + start, end - start, queries, &cfg, trace)) {
      all_ok = 0;
      break;
    }
    start = end;
  }
  for (U32 i = 0; i < count; i++)
    if (rep_of[i] != ~0u) qe_batch_copy_result(&queries[rep_of[i]], &queries[i]);
  for (U32 i = 0; i < count; i++)
    if (!queries[i].ok) all_ok = 0;
  free(rep_of);
  free(planned);
  if (trace) {
    U32 fused = 0;
    for (U32 i = 0; i < count; i++) fused += queries[i].fused ? 1u : 0u;
    (void)fallback;
    trace->queries = count;
    trace->fallback = count - fused;
    trace->fused = fused;
    QE_Trace tmp;
    qe_trace_end(&tmp, t0, p0);
    trace->gpu_kernel_ns = tmp.gpu_kernel_ns;
    trace->wall_ns = tmp.wall_ns;
  }
  return all_ok;
}

/* ---------------- autotuner ---------------- */
/* A batch's best configuration depends on the table's size, how wide the
 * union of its columns is and how many queries it has.  The tuner times
 * the candidate configurations on the caller's batch (a warm run, then a
 * timed run, wall clock: what a caller waits for), keeps the fastest per
 * workload key in memory and appends it to a text file
 * ($CQLHIP_TUNE_FILE, else ~/.cache/cqlhip/batch_tune.txt) that later
 * processes read back. */
typedef struct QE_TuneEntry {
  U32 rows_log2, width, count_log2;
  QE_BatchConfig cfg;
} QE_TuneEntry;

#define QE_TUNE_CAP 256
static QE_TuneEntry g_tune[QE_TUNE_CAP];
static U32 g_tune_count;
static B32 g_tune_loaded;
static pthread_mutex_t g_tune_lock = PTHREAD_MUTEX_INITIALIZER;

static U32 qe_log2_ceil(U64 x)
{
  U32 l = 0;
  while (((U64)1 << l) < x && l < 63) l++;
  return l;
}

static void qe_tune_path(char* out, size_t n)
{
  const char* env = getenv("CQLHIP_TUNE_FILE");
  if (env && env[0]) {
    snprintf(out, n, "%s", env);
    return;
  }
  const char* home = getenv("HOME");
  snprintf(out, n, "%s/.cache/cqlhip/batch_tune.txt", home ? home : "/tmp");
}

/* the workload key of a batch: rows, the byte width of the union of the
 * fusable queries' columns (per row), and the query count */
static void qe_tune_key(QE_Table* t, QE_BatchQuery* q, U32 count, U32* rows_log2, U32* width, U32* count_log2)
{
  QE_BatchStage st;
  memset(&st, 0, sizeof(st));
  QE_BatchPlanned p;
  for (U32 i = 0; i < count; i++)
    if (qe_batch_plan_query(t, &q[i], &p)) qe_batch_stage_query(t, &st, &p);
  U64 w = 0;
  for (U32 i = 0; i < st.count; i++) w += t->columns[st.column[i]].width;
  *rows_log2 = qe_log2_ceil(t->row_count ? t->row_count : 1);
  *width = (U32)((w + 15) / 16);
  *count_log2 = qe_log2_ceil(count ? count : 1);
}

static void qe_tune_load_locked(void)
{
  if (g_tune_loaded) return;
  g_tune_loaded = 1;
  char path[512];
  qe_tune_path(path, sizeof(path));
  FILE* f = fopen(path, "r");
  if (!f) return;
  QE_TuneEntry e;
  unsigned long long budget;
  while (g_tune_count < QE_TUNE_CAP &&
         fscanf(f, "%u %u %u %u %u %u %u %llu", &e.rows_log2, &e.width, &e.count_log2, &e.cfg.wg, &e.cfg.items,
                &e.cfg.buffers, &e.cfg.blocks_per_cu, &budget) == 8) {
    e.cfg.lds_budget = budget;
    g_tune[g_tune_count++] = e;
  }
  fclose(f);
}

static B32 qe_tune_lookup(U32 rl, U32 w, U32 cl, QE_BatchConfig* out)
{
  pthread_mutex_lock(&g_tune_lock);
  qe_tune_load_locked();
  B32 found = 0;
  for (U32 i = g_tune_count; i-- > 0;) {
    if (g_tune[i].rows_log2 == rl && g_tune[i].width == w && g_tune[i].count_log2 == cl) {
      *out = g_tune[i].cfg;
      found = 1;
      break;
    }
  }
  pthread_mutex_unlock(&g_tune_lock);
  return found;
}

static void qe_tune_store(U32 rl, U32 w, U32 cl, const QE_BatchConfig* cfg)
{
  pthread_mutex_lock(&g_tune_lock);
  qe_tune_load_locked();
  QE_TuneEntry e = {rl, w, cl, *cfg};
  if (g_tune_count < QE_TUNE_CAP) g_tune[g_tune_count++] = e;
  char path[512];
  qe_tune_path(path, sizeof(path));
  char dir[512];
  snprintf(dir, sizeof(dir), "%s", path);
  char* slash = strrchr(dir, '/');
  if (slash) {
    /* create the directory chain one level at a time */
    *slash = 0;
    for (char* p = dir + 1; *p; p++) {
      if (*p == '/') {
        *p = 0;
        mkdir(dir, 0755);
        *p = '/';
      }
    }
    mkdir(dir, 0755);
  }
  FILE* f = fopen(path, "a");
  if (f) {
    fprintf(f, "%u %u %u %u %u %u %u %llu\n", rl, w, cl, cfg->wg, cfg->items, cfg->buffers, cfg->blocks_per_cu,
            (unsigned long long)cfg->lds_budget);
    fclose(f);
  }
  pthread_mutex_unlock(&g_tune_lock);
}

B32 qe_aggregate_batch_tune(QE_Table* table, QE_BatchQuery* queries, U32 count, QE_BatchConfig* best, F64* best_ms)
{
  static const U32 shapes[][2] = {{1024, 1}, {512, 2}, {512, 1}, {256, 4}, {256, 1}};
  static const U32 buffers[] = {0, 2};
  static const U64 budgets[] = {64 * 1024, 0};
  F64 best_t = 1e300;
  QE_BatchConfig b;
  memset(&b, 0, sizeof(b));
  if (!qe_require_device(table, "qe_aggregate_batch_tune")) return 0;
  for (U32 i = 0; i < sizeof(shapes) / sizeof(shapes[0]); i++)
    for (U32 j = 0; j < 2; j++)
      for (U32 k = 0; k < 2; k++) {
        QE_BatchConfig c = {shapes[i][0], shapes[i][1], buffers[j], 2, budgets[k]};
        if (!qe_aggregate_batch_run(table, queries, count, &c, 0)) return 0;
        U64 t0 = qe_now_ns();
        if (!qe_aggregate_batch_run(table, queries, count, &c, 0)) return 0;
        F64 ms = (F64)(qe_now_ns() - t0) / 1e6;
        if (ms < best_t) {
          best_t = ms;
          b = c;
        }
      }
  U32 rl, w, cl;
  qe_tune_key(table, queries, count, &rl, &w, &cl);
  qe_tune_store(rl, w, cl, &b);
  if (best) *best = b;
  if (best_ms) *best_ms = best_t;
  return 1;
}

B32 qe_aggregate_batch(QE_Table* table, QE_BatchQuery* queries, U32 count, QE_BatchTrace* trace)
{
  QE_BatchConfig cfg;
  if (!qe_require_device(table, "qe_aggregate_batch")) return 0;
  qe_batch_config_from_settings(&cfg);
  if (table && queries && count && qe_settings_u64("QE_BATCH_AUTOTUNE", 0)) {
    /* the tuned configuration of this workload, tuning it on this batch the
     * first time */
    U32 rl, w, cl;
    qe_tune_key(table, queries, count, &rl, &w, &cl);
    if (!qe_tune_lookup(rl, w, cl, &cfg) && !qe_aggregate_batch_tune(table, queries, count, &cfg, 0)) return 0;
  }
  B32 ok = qe_aggregate_batch_run(table, queries, count, &cfg, trace);
  if (trace) {
    trace->tuned_wg = cfg.wg;
    trace->tuned_buffers = cfg.buffers;
  }
  return ok;
}
