/*
 * qe_join.c - aggregate over fact JOIN dim, pushed down to the fact scan.
 *
 *   qe_aggregate_join          qe_aggregate_join_pushdown, query_exec.c:8442-8775
 *   dense per-key reduction    qe_dense_stream_core :7632-7980 (small key range)
 *   hash probe                 qe_join_probe_stream_core :8080-8422 with
 *                              join_probe_native.comp
 *
 * The fact table is reduced on the GPU.  Three forms, chosen like the
 * reference chooses (dense first, else probe):
 *  - JoinDense: fact key range <= QE_AGG_DENSE_MAX_RANGE: totals per key
 *    value, then one host pass over the (filtered) dim rows adds each dim
 *    row's key totals into its group (query_exec.c:8621-8719);
 *  - JoinProbeGroup: the hash table's payload is the dim row's group code,
 *    so the probe aggregates straight into LDS-private group tables and no
 *    per-dim-row totals exist (MI355X form: one HBM pass, no scattered
 *    global atomics);
 *  - JoinProbeRow: payload = dim row, totals per dim row in HBM with global
 *    atomics, then the host combine (the reference's form).
 * Sums are exact 64-bit integers per key/group, converted to F64 and added
 * in F64 per dim row as the reference's combine does.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>

#include "qe_internal.h"

/* qe_aggregate_join_batch keeps the fact-side work one join computed for
 * the next join of the batch that needs the same: the fact rows in
 * partition order (fact, key column, argument columns, partition bits and
 * block ranges alike) and the dense per-key totals (fact, key column and
 * range, arguments, sums and min/max alike).  Outside a batch no work is
 * kept. */
typedef struct QE_JoinShare {
  B32 part_valid;
  const QE_Table* part_fact;
  U32 part_key_col, part_pbits, part_arg_count, part_B;
  U32 part_args[4];
  GPU_Buffer* hist;
  GPU_Buffer* off;
  GPU_Buffer* totals;
  GPU_Buffer* rows;
  B32 dense_valid;
  const QE_Table* dense_fact;
  U32 dense_key_col, dense_arg_count, dense_range, dense_kmin;
  U32 dense_args[4];
  B32 dense_sum, dense_mm;
  QE_DenseTotals dense_tot;
  U32 fact_passes_shared; /* joins that reused a kept fact side */
} QE_JoinShare;

static _Thread_local QE_JoinShare* t_join_share;

static void qe_join_share_drop_part(QE_JoinShare* sh)
{
  if (!sh->part_valid) return;
  qe_scratch_release(sh->hist);
  qe_scratch_release(sh->off);
  qe_scratch_release(sh->totals);
  qe_scratch_release(sh->rows);
  sh->hist = sh->off = sh->totals = sh->rows = 0;
  sh->part_valid = 0;
}

static void qe_join_share_drop_dense(QE_JoinShare* sh)
{
  if (!sh->dense_valid) return;
  qe_dense_totals_free(&sh->dense_tot);
  sh->dense_valid = 0;
}

typedef struct QE_GroupAcc {
  U32 value[QE_AGG_MAX_GROUP_COLS]; /* the dim group columns' codes (I32 biased), 0 past K */
  U64 count;
  F64 sum[QE_AGG_MAX_EXPRS];
  F64 mn[QE_AGG_MAX_EXPRS];
  F64 mx[QE_AGG_MAX_EXPRS];
  B32 seen;
} QE_GroupAcc;

/* ascending group values, column 0 first (codes of I32 columns are
 * biased, so this is signed order for them) */
static int qe_cmp_group(const void* a, const void* b)
{
  for (U32 k = 0; k < QE_AGG_MAX_GROUP_COLS; k++) {
    U32 x = ((const QE_GroupAcc*)a)->value[k], y = ((const QE_GroupAcc*)b)->value[k];
    if (x != y) return x < y ? -1 : 1;
  }
  return 0;
}

/* find-or-add group by value (open addressing over a pow2 table) */
typedef struct QE_GroupMap {
  U32* slots; /* group index + 1 */
  U64 mask;
  QE_GroupAcc* groups;
  U64 count;
} QE_GroupMap;

static B32 qe_group_map_init(QE_GroupMap* m, U64 max_groups)
{
  U64 size = 64;
  while (size < max_groups * 2 + 2) size <<= 1;
  m->slots = (U32*)calloc(size, sizeof(U32));
  m->groups = (QE_GroupAcc*)calloc(max_groups ? max_groups : 1, sizeof(QE_GroupAcc));
  m->mask = size - 1;
  m->count = 0;
  return m->slots && m->groups;
}

static QE_GroupAcc* qe_group_get(QE_GroupMap* m, const U32* value)
{
  U64 h = 0;
  for (U32 k = 0; k < QE_AGG_MAX_GROUP_COLS; k++) h = (h ^ value[k]) * 0x9E3779B97F4A7C15ull;
  h >>= 29;
  for (U64 s = h & m->mask;; s = (s + 1) & m->mask) {
    U32 g = m->slots[s];
    if (!g) {
      QE_GroupAcc* a = &m->groups[m->count++];
      memcpy(a->value, value, sizeof(a->value));
      m->slots[s] = (U32)m->count;
      return a;
    }
    if (!memcmp(m->groups[g - 1].value, value, QE_AGG_MAX_GROUP_COLS * sizeof(U32))) return &m->groups[g - 1];
  }
}

static void qe_group_map_free(QE_GroupMap* m)
{
  free(m->slots);
  free(m->groups);
}

/* add `matches` fact rows with the given key/row totals into a group,
 * query_exec.c:8693-8718 */
static void qe_group_add(QE_GroupAcc* g, const U32* funcs, const S32* expr_arg, U32 expr_count, const U32* arg_types,
                         U64 matches, const U64* sums, const U32* mins, const U32* maxs, U64 stride, U64 index)
{
  g->count += matches;
  for (U32 e = 0; e < expr_count; e++) {
    S32 a = expr_arg[e];
    if (funcs[e] == QE_AGG_FUNC_COUNT || a < 0) continue;
    U32 ty = arg_types[a];
    U64 s = sums ? sums[(U64)a * stride + index] : 0;
    F64 sv = ty == CQL_ColumnType_I32 ? (F64)(S64)s : (F64)s;
    U32 lo = mins ? mins[(U64)a * stride + index] : 0, hi = maxs ? maxs[(U64)a * stride + index] : 0;
    F64 mn = ty == CQL_ColumnType_I32 ? (F64)(S32)(lo ^ 0x80000000u) : (F64)lo;
    F64 mx = ty == CQL_ColumnType_I32 ? (F64)(S32)(hi ^ 0x80000000u) : (F64)hi;
    g->sum[e] += sv;
    g->mn[e] = g->seen ? fmin(g->mn[e], mn) : mn;
    g->mx[e] = g->seen ? fmax(g->mx[e], mx) : mx;
  }
  g->seen = 1;
}

static B32 qe_dim_u32(QE_Table* dim, U32 col, U32* out)
{
  if (dim->columns[col].type != CQL_ColumnType_U32 && dim->columns[col].type != CQL_ColumnType_I32) {
    qe_set_error("join: dim key/group columns must be U32/I32");
    return 0;
  }
  return qe_table_column_read(dim, col, 0, dim->row_count, out);
}

/* the radix-partitioned probe for large dimension tables: build per-
 * partition tables, scatter the fact rows into partition order, probe each
 * partition from LDS.  Fills host_acc (layout of qe_dense_totals_bytes) and
 * fl[2] (duplicate key, table full). */
static B32 qe_join_partitioned(QE_Table* fact, U32 fact_key_col, const U32* arg_col, const U32* arg_type, U32 arg_count,
                               B32 need_sum, B32 need_minmax, QE_Table* dim, U32 dim_key_col, GPU_Buffer* group_buf,
                               B32 group_mode, U32 gmin, B32 group_i32, GPU_Buffer* list, U64 pass_count, U64 R,
                               GPU_Buffer* acc, U64 acc_bytes, void* host_acc, U32* fl, QE_Trace* trace)
{
  /* 8192-slot tables (32 KB of key words + 32 KB of payloads in LDS), at
   * most half full; at least two partitions, so a slot's key word (the
   * hash bits below the partition bits) has its top bit free for the
   * occupied mark */
  const U32 sbits = 13, S = 1u << sbits;
  /* in a batch the tables are sized for the whole dim, so joins whose dim
   * WHEREs keep different row counts still partition the fact alike */
  QE_JoinShare* sh = t_join_share;
  const U64 size_keys = sh && dim->row_count > pass_count ? dim->row_count : pass_count;
  /* the packed table (group codes below 8192 beside the key bits in one
   * word, QE_JOIN_PROBE_PACKED, default 1) for the group-mode shapes that
   * have a packed probe; its key field needs at least 64 partitions */
  const B32 specialized = qe_settings_u64("QE_JOIN_PROBE_SPECIALIZED", 1) && qe_settings_u64("QE_GENERATE_JOIN", 1);
  const U32 pwg = qe_settings_u64("QE_JOIN_PROBE_WG", 512) == 256 ? 256 : 512;
  const B32 pk_shape = arg_count == 0 || (arg_count == 1 && (need_sum || need_minmax)) || (arg_count == 2 && need_sum);
  const B32 packed = group_mode && specialized && pwg == 512 && pk_shape && R <= 8192 &&
                     qe_settings_u64("QE_JOIN_PROBE_PACKED", 1) != 0;
  U32 pbits = packed ? 6 : 1;
  while (((U64)1 << pbits) * (S / 2) < size_keys && pbits < 12) pbits++;
  const U64 P = 1ull << pbits;
  if (P * (S * 7 / 10) < pass_count) {
    qe_set_error("qe_aggregate_join: %llu dim keys exceed the partitioned tables", (unsigned long long)pass_count);
    return 0;
  }
  const U64 rows = fact->row_count;
  const U32 W = 1 + arg_count;
  /* rows of 1 or 2 words go through join_part_scatter_wc (whole 128-byte
   * lines, partial ones carried in LDS) when two of its 512-thread blocks
   * fit a CU: next u64[P] | 4 u32[P] | lines 16 B[P + 4096 / CH] | carry
   * 128 B[P] | rows u32[4096 * W], CH = 32 / W */
  const U64 wc_lds = P * 24 + 16 * (P + 4096 / (32 / (W > 2 ? 1 : W))) + P * 128 + 4096 * 4 * (U64)W;
  const B32 wc = W <= 2 && rows < (1ull << 32) && qe_settings_u64("QE_JOIN_SCATTER_LINES", 1) &&
                 2 * (wc_lds + 64) <= gpu_hip_max_lds_bytes();
  const U32 swg = wc ? 512 : QE_WG;
  /* hist and scatter blocks (the same block ranges): the scatter's blocks
   * per CU are bounded by its LDS, separately from the scan kernels' */
  U32 B = qe_stream_blocks((rows + 3) / 4);
  {
    U64 cap = gpu_hip_compute_units() * (wc ? qe_settings_u64("QE_JOIN_SCATTER_WC_BLOCKS_PER_CU", 2)
                                            : qe_settings_u64("QE_JOIN_SCATTER_BLOCKS_PER_CU", 4));
    if (cap >= 1 && B > cap) B = (U32)cap;
  }
  GPU_Kernel* kb = qe_kernel("join_part_build");
  /* the histogram runs with the scatter's block count and block size */
  GPU_Kernel* kh = qe_kernel(wc ? "join_part_hist_512" : "join_part_hist");
  GPU_Kernel* ko = qe_kernel("join_part_offsets");
  /* the instances specialised to the query's shape (`specialized` above)
   * stand in for the reference's generated join kernel (QE_GENERATE_JOIN,
   * query_exec.c:8235); 0 takes the generic forms */
  char sname[40] = "join_part_scatter";
  if (wc) snprintf(sname, sizeof(sname), "join_part_scatter_wc_%u", arg_count);
  else if (arg_count <= 2 && specialized)
    snprintf(sname, sizeof(sname), "join_part_scatter_%u", arg_count);
  GPU_Kernel* ks = qe_kernel(sname);
  /* a probe instance specialised for this shape when one exists */
  char pname[48] = "join_part_probe";
  if (group_mode) {
    if (arg_count == 0) snprintf(pname, sizeof(pname), "join_part_probe_0");
    else if (arg_count <= 2 && need_sum) snprintf(pname, sizeof(pname), "join_part_probe_%u%s", arg_count, need_minmax ? "sm" : "s");
    else if (arg_count == 1 && need_minmax) snprintf(pname, sizeof(pname), "join_part_probe_1m");
  } else {
    if (arg_count == 0) snprintf(pname, sizeof(pname), "join_part_probe_0_row");
    else if (arg_count == 1 && need_sum) snprintf(pname, sizeof(pname), "join_part_probe_1%s_row", need_minmax ? "sm" : "s");
  }
  if (!specialized) snprintf(pname, sizeof(pname), "join_part_probe");
  /* QE_JOIN_PROBE_WG: 512 (default) or 256 threads a probe block */
  if (packed) {
    const char* shape = arg_count == 0 ? "0" : (arg_count == 1 ? (need_sum ? (need_minmax ? "1sm" : "1s") : "1m")
                                                               : (need_minmax ? "2sm" : "2s"));
    snprintf(pname, sizeof(pname), "join_part_probe_pk_%s", shape);
  }
  if (pwg == 512) snprintf(pname + strlen(pname), sizeof(pname) - strlen(pname), "_512");
  GPU_Kernel* kp = qe_kernel(pname);
  GPU_Buffer* table = qe_scratch("jp_table", P * S * 8);
  GPU_Buffer* flags = qe_scratch("join_flags", 64);
  /* the fact rows in partition order that an earlier join of the batch
   * left: the same fact, key, arguments, partitions and block ranges */
  const B32 reuse = sh && sh->part_valid && sh->part_fact == fact && sh->part_key_col == fact_key_col &&
                    sh->part_pbits == pbits && sh->part_B == B && sh->part_arg_count == arg_count &&
                    memcmp(sh->part_args, arg_col, arg_count * sizeof(U32)) == 0;
  if (sh && !reuse) qe_join_share_drop_part(sh);
  GPU_Buffer* hist = reuse ? sh->hist : qe_scratch("jp_hist", (U64)B * P * 4);
  GPU_Buffer* off = reuse ? sh->off : qe_scratch("jp_off", (U64)B * P * 4);
  GPU_Buffer* totals = reuse ? sh->totals : qe_scratch("jp_totals", P * 4);
  /* rows past the partitions' rows: the last takes the scatter's discarded
   * stores, and the packed probe's 16-byte loads of 2 or 4 rows may reach
   * up to 3 rows past the partitions' last row */
  GPU_Buffer* rowsbuf = reuse ? sh->rows : qe_scratch("jp_rows", (rows + 4) * W * 4);
  B32 ok = kb && kh && ko && ks && kp && table && flags && hist && off && totals && rowsbuf;
  U64 cus = gpu_hip_compute_units();
  U32 split = (U32)((cus * qe_settings_u64("QE_JOIN_PROBE_BLOCKS_PER_CU", 24) + P - 1) / P);
  if (split < 1) split = 1;
  /* key words u32[S], then payloads: u16 group codes or u32 dim rows; in
   * group mode 2^clog copies of the group tables (cnt u32[R], sums u64[E][R],
   * min/max u32[E][R], each with an odd stride).  The copies are as many as
   * QE_JOIN_PROBE_COPIES (default 2) allows while QE_JOIN_PROBE_BLOCKS_PER_CU_LDS
   * (default 3) probe blocks still fit a CU's LDS, at least one. */
  const U64 table_lds = (U64)S * 4 + (packed ? 0 : (U64)S * (group_mode ? 2 : 4));
  U32 clog = 0;
  U64 acc_lds = 0;
  if (group_mode) {
    const U64 ew = R * arg_count;
    const U64 copy_bytes = (R | 1) * 4 + (need_sum && ew ? (ew | 1) * 8 : 0) + (need_minmax && ew ? 2 * (ew | 1) * 4 : 0);
    const U64 max_copies = qe_settings_u64("QE_JOIN_PROBE_COPIES", 2);
    U64 fit = qe_settings_u64("QE_JOIN_PROBE_BLOCKS_PER_CU_LDS", 3);
    const U64 per_block = gpu_hip_kernel_max_lds_bytes(kp) / (fit ? fit : 1);
    while (clog < 4 && (2ull << clog) <= max_copies && table_lds + (2ull << clog) * copy_bytes + 64 <= per_block) clog++;
    acc_lds = ((U64)1 << clog) * copy_bytes;
  }
  U64 probe_lds = table_lds + acc_lds;
  if (ok && probe_lds > gpu_hip_kernel_max_lds_bytes(kp)) {
    qe_set_error("qe_aggregate_join: partition probe needs %llu bytes of LDS", (unsigned long long)probe_lds);
    ok = 0;
  }
  /* scatter: next, dst u64[P] | cnt, start u32[P] | rows u32[4096 * W], plus
   * its 16 static bytes */
  const U64 scatter_lds = wc ? wc_lds : P * 24 + 4096 * 4 * (U64)W;
  if (ok && scatter_lds + 64 > gpu_hip_max_lds_bytes()) {
    qe_set_error("qe_aggregate_join: partition scatter needs %llu bytes of LDS", (unsigned long long)scatter_lds);
    ok = 0;
  }
  if (ok) {
    /* build */
    gpu_kernel_set_arg_buffer(kb, 0, dim->columns[dim_key_col].buffer);
    if (group_buf) gpu_kernel_set_arg_buffer(kb, 1, group_buf);
    else gpu_hip_kernel_clear_arg(kb, 1);
    gpu_kernel_set_arg_buffer(kb, 2, table);
    gpu_kernel_set_arg_buffer(kb, 3, flags);
    if (list) gpu_kernel_set_arg_buffer(kb, 4, list);
    else gpu_hip_kernel_clear_arg(kb, 4);
    gpu_kernel_set_arg_u64(kb, 0, pass_count);
    qe_kernel_set_word(kb, QK_W_PMODE, (group_mode ? (group_buf ? 1u : 2u) : 0u) | (packed ? 256u : 0u));
    qe_kernel_set_word(kb, QK_W_GMIN, gmin);
    qe_kernel_set_word(kb, QK_W_GTYPE, group_i32 ? CQL_ColumnType_I32 : CQL_ColumnType_U32);
    qe_kernel_set_word(kb, QK_W_PBITS, pbits);
    qe_kernel_set_word(kb, QK_W_SBITS, sbits);
    /* fact columns: slot 0 the key, then the args */
    QE_ColumnSlots slots;
    memset(&slots, 0, sizeof(slots));
    qe_slot_of(&slots, fact_key_col);
    U32 slot_bits = 0, type_bits = 0;
    for (U32 a = 0; a < arg_count; a++) {
      slot_bits |= qe_slot_of(&slots, arg_col[a]) << (4 * a);
      type_bits |= arg_type[a] << (8 * a);
    }
    GPU_Kernel* fk_kernels[2] = {kh, ks};
    for (U32 i = 0; i < 2; i++) {
      GPU_Kernel* k = fk_kernels[i];
      for (U32 sl = 0; sl < QE_MAX_COLUMN_BINDINGS; sl++) {
        if (sl < slots.count) gpu_kernel_set_arg_buffer(k, 6 + sl, fact->columns[slots.column[sl]].buffer);
        else gpu_hip_kernel_clear_arg(k, 6 + sl);
      }
      gpu_kernel_set_arg_u64(k, 0, rows);
      gpu_kernel_set_arg_u64(k, 4, 0);
      gpu_kernel_set_arg_u64(k, 5, 0);
      qe_kernel_set_word(k, QK_W_PBITS, pbits);
      qe_kernel_set_word(k, QK_W_ROWW, W);
      qe_kernel_set_word(k, QK_W_ARGS, arg_count | (slot_bits << 8));
    }
    gpu_kernel_set_arg_buffer(kh, 2, hist);
    gpu_hip_kernel_set_lds_bytes(kh, (U32)(P * 4));
    gpu_kernel_set_arg_buffer(ko, 2, hist);
    gpu_kernel_set_arg_buffer(ko, 3, off);
    gpu_kernel_set_arg_buffer(ko, 4, totals);
    gpu_kernel_set_arg_u64(ko, 0, B);
    qe_kernel_set_word(ko, QK_W_PBITS, pbits);
    gpu_kernel_set_arg_buffer(ks, 2, rowsbuf);
    gpu_kernel_set_arg_buffer(ks, 3, off);
    gpu_kernel_set_arg_buffer(ks, 4, totals);
    gpu_hip_kernel_set_lds_bytes(ks, (U32)scatter_lds);
    U32 kflags = (need_sum ? 1u : 0u) | (need_minmax ? 2u : 0u) | (group_mode ? 0u : 4u);
    gpu_kernel_set_arg_buffer(kp, 1, acc);
    gpu_kernel_set_arg_buffer(kp, 2, table);
    gpu_kernel_set_arg_buffer(kp, 3, rowsbuf);
    gpu_kernel_set_arg_buffer(kp, 4, totals);
    qe_kernel_set_word(kp, QK_W_PBITS, pbits);
    qe_kernel_set_word(kp, QK_W_SBITS, sbits);
    qe_kernel_set_word(kp, QK_W_ROWW, W);
    qe_kernel_set_word(kp, QK_W_SPLIT, split);
    qe_kernel_set_word(kp, QK_W_ARGS, arg_count | (clog << 16));
    qe_kernel_set_word(kp, QK_W_TYPES, type_bits);
    qe_kernel_set_word(kp, QK_W_RANGE, (U32)R);
    qe_kernel_set_word(kp, QK_W_KEY, kflags << 16);
    {
      /* the packed probe's count-and-sum atomic: the argument's largest
       * value (zone maps), all of U32 when they are not built */
      U32 amax = 0xffffffffu;
      if (arg_count == 1 && arg_type[0] == CQL_ColumnType_U32 && fact->columns[arg_col[0]].zonemap_valid &&
          fact->columns[arg_col[0]].max_value >= 0.0 && fact->columns[arg_col[0]].max_value < 4294967295.0)
        amax = (U32)fact->columns[arg_col[0]].max_value;
      qe_kernel_set_word(kp, QK_W_GMIN, amax);
    }
    gpu_hip_kernel_set_lds_bytes(kp, (U32)probe_lds);
    U64 build_threads = pass_count < 256 ? 256 : pass_count;
    if (build_threads > cus * 4 * 256) build_threads = cus * 4 * 256;
    U64 zero_bytes = R * 8 * (1 + arg_count);
    GPU_Batch* b = gpu_batch_begin(0, acc_bytes + 8);
    gpu_batch_buffer_zero(b, table, P * S * 8);
    gpu_batch_buffer_zero(b, flags, 8);
    gpu_batch_buffer_zero(b, acc, acc_bytes);
    if (arg_count && need_minmax) {
      gpu_batch_buffer_fill(b, acc, zero_bytes + R * 4 * arg_count, 0xffffffffu);
      gpu_batch_buffer_zero(b, acc, zero_bytes);
    }
    gpu_batch_kernel_execute(b, kb, (U32)build_threads, QE_WG);
    if (!reuse) {
      gpu_batch_kernel_execute(b, kh, B * swg, swg);
      gpu_batch_kernel_execute(b, ko, (U32)(P * QE_WG), QE_WG);
      gpu_batch_kernel_execute(b, ks, B * swg, swg);
    }
    gpu_batch_kernel_execute(b, kp, (U32)(P * split * pwg), pwg);
    gpu_batch_buffer_read(b, flags, fl, 8);
    gpu_batch_buffer_read(b, acc, host_acc, acc_bytes);
    ok = gpu_batch_end(b);
    gpu_hip_kernel_set_lds_bytes(kh, 0);
    gpu_hip_kernel_set_lds_bytes(ks, 0);
    gpu_hip_kernel_set_lds_bytes(kp, 0);
    if (!ok) qe_set_error("partitioned join: %s", gpu_hip_last_error());
    if (trace) {
      trace->kernel_count = reuse ? 2 : 5;
      trace->rows_scanned = rows;
      trace->bytes_read = reuse ? rows * 4 * (U64)W : rows * 4 * (U64)slots.count;
      snprintf(trace->kernel, sizeof(trace->kernel), "%s", pname);
    }
  }
  qe_scratch_release(table);
  qe_scratch_release(flags);
  if (sh && ok) {
    /* kept for the next join of the batch */
    if (reuse) {
      sh->fact_passes_shared++;
    } else {
      sh->part_valid = 1;
      sh->part_fact = fact;
      sh->part_key_col = fact_key_col;
      sh->part_pbits = pbits;
      sh->part_B = B;
      sh->part_arg_count = arg_count;
      memcpy(sh->part_args, arg_col, arg_count * sizeof(U32));
      sh->hist = hist;
      sh->off = off;
      sh->totals = totals;
      sh->rows = rowsbuf;
    }
  } else {
    if (sh && reuse) {
      sh->hist = sh->off = sh->totals = sh->rows = 0;
      sh->part_valid = 0;
    }
    qe_scratch_release(hist);
    qe_scratch_release(off);
    qe_scratch_release(totals);
    qe_scratch_release(rowsbuf);
  }
  return ok;
}

B32 qe_aggregate_join(QE_Table* fact, U32 fact_key_col, QE_Table* dim, U32 dim_key_col, const QE_IRNode* dim_where,
                      U32 dim_where_count, const S32* dim_group_cols, U32 dim_group_count, const U32* funcs,
                      const S32* fact_arg_cols, U32 expr_count, U64 cap, U64* out_group_count, F64* out_keys,
                      U64* out_cells, QE_Trace* trace)
{
  U64 t0, p0;
  qe_trace_begin(trace, &t0, &p0);
  if (!fact || !dim || !funcs || !fact_arg_cols || !out_group_count || !out_cells || fact_key_col >= fact->column_count ||
      dim_key_col >= dim->column_count || dim_group_count > QE_AGG_MAX_GROUP_COLS ||
      (dim_group_count && !dim_group_cols) || expr_count == 0 || expr_count > QE_AGG_MAX_EXPRS) {
    qe_set_error("qe_aggregate_join: bad arguments");
    return 0;
  }
  const U32 K = dim_group_count;
  for (U32 k = 0; k < K; k++)
    if (dim_group_cols[k] < 0 || (U32)dim_group_cols[k] >= dim->column_count) {
      qe_set_error("qe_aggregate_join: bad dim group column");
      return 0;
    }
  if (!qe_require_device(fact, "qe_aggregate_join") || !qe_require_device(dim, "qe_aggregate_join")) return 0;
  if (!qe_require_non_null(fact, (S32)fact_key_col, "qe_aggregate_join") ||
      !qe_require_non_null(dim, (S32)dim_key_col, "qe_aggregate_join") ||
      !qe_where_non_null(dim, dim_where, dim_where_count, "qe_aggregate_join"))
    return 0;
  for (U32 k = 0; k < K; k++)
    if (!qe_require_non_null(dim, dim_group_cols[k], "qe_aggregate_join")) return 0;
  for (U32 e = 0; e < expr_count; e++)
    if (!qe_require_non_null(fact, fact_arg_cols[e], "qe_aggregate_join")) return 0;
  /* the pushdown rests on the streamed reductions (query_exec.c:8446) */
  if (qe_settings_u64("QE_JOIN_PUSHDOWN", 1) == 0 || qe_settings_u64("QE_AGG_STREAM", 1) == 0) {
    qe_set_error("qe_aggregate_join: QE_JOIN_PUSHDOWN or QE_AGG_STREAM is off");
    return 0;
  }
  QE_Column* fkey = &fact->columns[fact_key_col];
  if (fkey->type != CQL_ColumnType_U32 && fkey->type != CQL_ColumnType_I32) {
    qe_set_error("qe_aggregate_join: fact key must be U32/I32");
    return 0;
  }
  /* fact-side aggregate arguments (distinct columns, <= 4) */
  U32 arg_col[4], arg_type[4], arg_count = 0;
  S32 expr_arg[QE_AGG_MAX_EXPRS];
  B32 need_sum = 0, need_minmax = 0;
  for (U32 e = 0; e < expr_count; e++) {
    expr_arg[e] = -1;
    if (funcs[e] > QE_AGG_FUNC_MAX) {
      qe_set_error("qe_aggregate_join: bad function");
      return 0;
    }
    if (fact_arg_cols[e] < 0 || funcs[e] == QE_AGG_FUNC_COUNT) {
      if (fact_arg_cols[e] < 0 && funcs[e] != QE_AGG_FUNC_COUNT) {
        qe_set_error("qe_aggregate_join: function needs an argument");
        return 0;
      }
      continue;
    }
    U32 c = (U32)fact_arg_cols[e];
    if (c >= fact->column_count ||
        (fact->columns[c].type != CQL_ColumnType_U32 && fact->columns[c].type != CQL_ColumnType_I32)) {
      qe_set_error("qe_aggregate_join: fact arguments must be U32/I32 columns");
      return 0;
    }
    U32 a = 0;
    while (a < arg_count && arg_col[a] != c) a++;
    if (a == arg_count) {
      if (arg_count == 4) {
        qe_set_error("qe_aggregate_join: more than 4 argument columns");
        return 0;
      }
      arg_col[arg_count] = c;
      arg_type[arg_count] = fact->columns[c].type;
      arg_count++;
    }
    expr_arg[e] = (S32)a;
    if (funcs[e] == QE_AGG_FUNC_SUM || funcs[e] == QE_AGG_FUNC_AVG) need_sum = 1;
    if (funcs[e] == QE_AGG_FUNC_MIN || funcs[e] == QE_AGG_FUNC_MAX) need_minmax = 1;
  }
  if (trace) trace->rows_before = fact->row_count;

  B32 key_i32 = fkey->type == CQL_ColumnType_I32;
  if (dim->columns[dim_key_col].type != CQL_ColumnType_U32 && dim->columns[dim_key_col].type != CQL_ColumnType_I32) {
    qe_set_error("qe_aggregate_join: the dim key must be U32/I32");
    return 0;
  }
  /* the dim group columns and the ranges their zone maps give */
  QE_GroupKeys gk;
  if (!qe_group_keys(dim, dim_group_cols, K, &gk)) return 0;
  B32 gi32[QE_AGG_MAX_GROUP_COLS];
  for (U32 k = 0; k < K; k++) {
    gi32[k] = gk.type[k] == CQL_ColumnType_I32;
    if (gk.type[k] != CQL_ColumnType_U32 && !gi32[k]) {
      qe_set_error("qe_aggregate_join: dim group columns must be U32/I32");
      return 0;
    }
  }
  U64 dn = dim->row_count;
  F64 kspan = fkey->max_value - fkey->min_value + 1.0;
  B32 dense = fact->row_count > 0 && qe_settings_u64("QE_JOIN_DENSE", 1) != 0 &&
              kspan <= (F64)qe_settings_u64("QE_AGG_DENSE_MAX_RANGE", 65536);
  /* group codes from the dim group columns' zone maps (a superset of the
   * passing rows' range); several columns combine as digits */
  U32 gmin = K == 1 ? gk.min_code[0] : 0u;
  U64 grange = K == 0 ? 1 : qe_group_keys_span(&gk);
  if (!dn) grange = 1;
  U64 lds_need = grange * 4 + (need_sum ? grange * 8 * arg_count : 0) + (need_minmax ? grange * 8 * arg_count : 0);
  B32 group_mode = !dense && qe_settings_u64("QE_JOIN_GROUP_PAYLOAD", 1) != 0 && grange != 0 &&
                   grange <= qe_settings_u64("QE_JOIN_GROUP_MAX", 8192) && lds_need <= gpu_hip_max_lds_bytes();
  if (!dense && qe_settings_u64("QE_JOIN_PROBE", 1) == 0) {
    qe_set_error("qe_aggregate_join: key range too wide for the dense form and QE_JOIN_PROBE is off");
    return 0;
  }

  /* dim rows passing the dim WHERE: on the device for the probe forms, on
   * the host for the forms that combine per dim row */
  GPU_Buffer* pass_dev = 0;
  U64 pass_count = dn;
  U64* pass = 0;
  U32* dkeys = 0;
  U32* dgroups[QE_AGG_MAX_GROUP_COLS] = {0, 0, 0, 0};
  GPU_Buffer* code_buf = 0; /* several group columns in group mode: each dim row's digits */
  B32 ok = 1;
  B32 host_dim = dense || !group_mode;
  if (dim_where && dim_where_count) {
    if (host_dim) {
      pass = (U64*)malloc(sizeof(U64) * (dn ? dn : 1));
      ok = pass && qe_scan_filter(dim, dim_where, dim_where_count, 0, pass, dn, &pass_count, 0);
    } else {
      ok = qe_scan_filter_device(dim, dim_where, dim_where_count, &pass_dev, &pass_count);
      if (ok && pass_dev && qe_settings_u64("QE_JOIN_VERIFY_LIST", 0)) {
        /* debug: the device selection must be strictly ascending row ids */
        U64* chk = (U64*)malloc(sizeof(U64) * (pass_count ? pass_count : 1));
        if (chk) {
          ok = qe_buffer_read_checked(pass_dev, chk, pass_count * 8);
          for (U64 i = 0; ok && i < pass_count; i++) {
            if (chk[i] >= dn || (i && chk[i] <= chk[i - 1])) {
              qe_set_error("qe_aggregate_join: device selection broken at %llu: %llu after %llu", (unsigned long long)i,
                           (unsigned long long)chk[i], (unsigned long long)(i ? chk[i - 1] : 0));
              ok = 0;
              break;
            }
          }
          free(chk);
        }
      }
    }
  } else if (host_dim) {
    pass = (U64*)malloc(sizeof(U64) * (dn ? dn : 1));
    ok = pass != 0;
    for (U64 i = 0; ok && i < dn; i++) pass[i] = i;
  }
  if (ok && host_dim) {
    dkeys = (U32*)malloc(sizeof(U32) * (dn ? dn : 1));
    ok = dkeys && qe_dim_u32(dim, dim_key_col, dkeys);
  }
  if (ok && (host_dim || K > 1)) {
    for (U32 k = 0; ok && k < K; k++) {
      dgroups[k] = (U32*)malloc(sizeof(U32) * (dn ? dn : 1));
      ok = dgroups[k] && qe_dim_u32(dim, (U32)dim_group_cols[k], dgroups[k]);
    }
  }
  if (ok && group_mode && K > 1) {
    /* the digits of each dim row's group values, key 0 most significant:
     * the build kernel takes them as a U32 group column with minimum 0 */
    U32* codes = (U32*)malloc(sizeof(U32) * (dn ? dn : 1));
    ok = codes != 0;
    for (U64 r = 0; ok && r < dn; r++) {
      U64 c = 0;
      for (U32 k = 0; k < K; k++) {
        const U32 v = gi32[k] ? dgroups[k][r] ^ 0x80000000u : dgroups[k][r];
        c = c * gk.range[k] + (U64)(v - gk.min_code[k]);
      }
      codes[r] = (U32)c;
    }
    if (ok) {
      code_buf = qe_scratch("join_group_codes", (dn ? dn : 1) * 4);
      ok = code_buf && qe_buffer_write_checked(code_buf, codes, dn * 4);
    }
    free(codes);
  }

  QE_GroupMap groups;
  memset(&groups, 0, sizeof(groups));
  if (ok) ok = qe_group_map_init(&groups, (host_dim ? pass_count : grange) + 1);
  U32 path = 0;
  QE_Prune all;
  memset(&all, 0, sizeof(all));
  all.row_begin = 0;
  all.row_end = fact->row_count;
  all.zones_total = all.zones_left = 1;
  /* a dim row's group tuple (codes, I32 biased) */
  U32 gv[QE_AGG_MAX_GROUP_COLS];
#define QE_DIM_GROUP(d)                                                                                                \
  do {                                                                                                                 \
    memset(gv, 0, sizeof(gv));                                                                                         \
    for (U32 k_ = 0; k_ < K; k_++) gv[k_] = gi32[k_] ? dgroups[k_][(d)] ^ 0x80000000u : dgroups[k_][(d)];              \
  } while (0)

  if (!ok || fact->row_count == 0 || pass_count == 0) {
    /* nothing to join */
  } else if (dense) {
    path = QE_AggPath_JoinDense;
    U32 range = (U32)kspan;
    QE_GroupKeys fk;
    memset(&fk, 0, sizeof(fk));
    fk.count = 1;
    fk.col[0] = fact_key_col;
    fk.type[0] = fkey->type;
    fk.min_code[0] = key_i32 ? ((U32)(S32)fkey->min_value ^ 0x80000000u) : (U32)fkey->min_value;
    fk.range[0] = range;
    const U32 kmin_code = fk.min_code[0];
    QE_DenseTotals tot;
    memset(&tot, 0, sizeof(tot));
    /* per-key totals an earlier join of the batch left: same fact, key and
     * range, arguments, sums and min/max */
    QE_JoinShare* sh = t_join_share;
    const B32 reuse = sh && sh->dense_valid && sh->dense_fact == fact && sh->dense_key_col == fact_key_col &&
                      sh->dense_range == range && sh->dense_kmin == kmin_code && sh->dense_arg_count == arg_count &&
                      sh->dense_sum == need_sum && sh->dense_mm == need_minmax &&
                      memcmp(sh->dense_args, arg_col, arg_count * sizeof(U32)) == 0;
    if (reuse) {
      tot = sh->dense_tot;
      sh->fact_passes_shared++;
    } else {
      if (sh) qe_join_share_drop_dense(sh);
      ok = qe_dense_run(fact, &fk, arg_col, arg_count, need_sum, need_minmax, 0, &all, &tot, trace);
      if (ok && sh) {
        sh->dense_valid = 1;
        sh->dense_fact = fact;
        sh->dense_key_col = fact_key_col;
        sh->dense_range = range;
        sh->dense_kmin = kmin_code;
        sh->dense_arg_count = arg_count;
        memcpy(sh->dense_args, arg_col, arg_count * sizeof(U32));
        sh->dense_sum = need_sum;
        sh->dense_mm = need_minmax;
        sh->dense_tot = tot;
      }
    }
    if (ok) {
      for (U64 v = 0; v < pass_count; v++) {
        U64 d = pass[v];
        U32 code = key_i32 ? (dkeys[d] ^ 0x80000000u) : dkeys[d];
        U32 kid = code - kmin_code;
        if (code < kmin_code || kid >= range || tot.count[kid] == 0) continue;
        QE_DIM_GROUP(d);
        QE_GroupAcc* g = qe_group_get(&groups, gv);
        qe_group_add(g, funcs, expr_arg, expr_count, arg_type, tot.count[kid], tot.sum, tot.mn, tot.mx, range, kid);
      }
      if (!sh) qe_dense_totals_free(&tot);
    }
  } else {
    path = group_mode ? QE_AggPath_JoinProbeGroup : QE_AggPath_JoinProbeRow;
    U64 R = group_mode ? grange : dn;
    U64 slots_n = 64;
    while (slots_n < pass_count * 2) slots_n <<= 1;
    U32 lg = 0;
    while ((1ull << lg) < slots_n) lg++;
    GPU_Kernel* kb = qe_kernel("join_build");
    GPU_Kernel* kp = qe_kernel("join_probe");
    U64 acc_bytes = qe_dense_totals_bytes((U32)R, arg_count);
    GPU_Buffer* table = qe_scratch("join_table", slots_n * 8);
    GPU_Buffer* flags = qe_scratch("join_flags", 64);
    GPU_Buffer* acc = qe_scratch("join_totals", acc_bytes);
    GPU_Buffer* list_buf = 0;
    if (!group_mode && pass_count != dn) {
      list_buf = qe_scratch("join_list", pass_count * 8);
      if (list_buf && !qe_buffer_write_checked(list_buf, pass, pass_count * 8)) {
        qe_scratch_release(list_buf);
        list_buf = 0;
      }
    }
    /* group mode's payload: the one group column, the digits of several,
     * or none (one group) */
    GPU_Buffer* group_buf = !group_mode ? 0 : (K == 1 ? dim->columns[dim_group_cols[0]].buffer : code_buf);
    const B32 g_i32 = K == 1 && gi32[0];
    void* host_acc = malloc(acc_bytes ? acc_bytes : 8);
    ok = kb && kp && table && flags && acc && host_acc && (group_mode || pass_count == dn || list_buf);
    U32 fl[2] = {0, 0};
    B32 partitioned = qe_settings_u64("QE_JOIN_PARTITION", 1) != 0 &&
                      pass_count >= qe_settings_u64("QE_JOIN_PARTITION_MIN_KEYS", 32768) && fact->row_count >= (1u << 20);
    if (ok && partitioned) {
      path = group_mode ? QE_AggPath_JoinPartGroup : QE_AggPath_JoinPartRow;
      ok = qe_join_partitioned(fact, fact_key_col, arg_col, arg_type, arg_count, need_sum, need_minmax, dim, dim_key_col,
                               group_buf, group_mode, gmin, g_i32, pass_dev ? pass_dev : list_buf, pass_count, R, acc,
                               acc_bytes, host_acc, fl, trace);
      if (ok && fl[0]) {
        ok = 0;
        qe_set_error("qe_aggregate_join: duplicate dim keys (the probe form needs unique keys)");
      }
      if (ok && fl[1]) {
        ok = 0;
        qe_set_error("qe_aggregate_join: partition table full");
      }
    } else if (ok) {
      gpu_kernel_set_arg_buffer(kb, 0, dim->columns[dim_key_col].buffer);
      if (group_buf) gpu_kernel_set_arg_buffer(kb, 1, group_buf);
      else gpu_hip_kernel_clear_arg(kb, 1);
      gpu_kernel_set_arg_buffer(kb, 2, table);
      gpu_kernel_set_arg_buffer(kb, 3, flags);
      if (pass_dev) gpu_kernel_set_arg_buffer(kb, 4, pass_dev);
      else if (list_buf) gpu_kernel_set_arg_buffer(kb, 4, list_buf);
      else gpu_hip_kernel_clear_arg(kb, 4);
      gpu_kernel_set_arg_u64(kb, 0, pass_count);
      qe_kernel_set_word(kb, QK_W_LOG2, lg);
      qe_kernel_set_word(kb, QK_W_PMODE, group_mode ? (group_buf ? 1u : 2u) : 0u);
      qe_kernel_set_word(kb, QK_W_GMIN, gmin);
      qe_kernel_set_word(kb, QK_W_GTYPE, g_i32 ? CQL_ColumnType_I32 : CQL_ColumnType_U32);
      QE_ColumnSlots slots;
      memset(&slots, 0, sizeof(slots));
      qe_slot_of(&slots, fact_key_col); /* slot 0: the fact key */
      U32 slot_bits = 0, type_bits = 0;
      for (U32 a = 0; a < arg_count; a++) {
        slot_bits |= qe_slot_of(&slots, arg_col[a]) << (4 * a);
        type_bits |= arg_type[a] << (8 * a);
      }
      for (U32 s = 0; s < QE_MAX_COLUMN_BINDINGS; s++) {
        if (s < slots.count) gpu_kernel_set_arg_buffer(kp, 6 + s, fact->columns[slots.column[s]].buffer);
        else gpu_hip_kernel_clear_arg(kp, 6 + s);
      }
      U32 kflags = (need_sum ? 1u : 0u) | (need_minmax ? 2u : 0u) | (group_mode ? 0u : 4u);
      gpu_kernel_set_arg_buffer(kp, 1, acc);
      gpu_kernel_set_arg_buffer(kp, 2, table);
      gpu_hip_kernel_clear_arg(kp, 5);
      gpu_kernel_set_arg_u64(kp, 0, fact->row_count);
      gpu_kernel_set_arg_u64(kp, 4, 0);
      gpu_kernel_set_arg_u64(kp, 5, 0);
      qe_kernel_set_word(kp, QK_W_LOG2, lg);
      qe_kernel_set_word(kp, QK_W_RANGE, (U32)R);
      qe_kernel_set_word(kp, QK_W_ARGS, arg_count | (slot_bits << 8));
      qe_kernel_set_word(kp, QK_W_TYPES, type_bits);
      qe_kernel_set_word(kp, QK_W_KEY, kflags << 16);
      gpu_hip_kernel_set_lds_bytes(kp, group_mode ? (U32)lds_need : 0u);
      U64 quads = (fact->row_count + 3) / 4;
      U32 blocks = qe_stream_blocks(quads);
      U64 build_threads = pass_count < 256 ? 256 : pass_count;
      if (build_threads > (U64)gpu_hip_compute_units() * 4 * 256) build_threads = (U64)gpu_hip_compute_units() * 4 * 256;
      U64 zero_bytes = R * 8 * (1 + arg_count);
      GPU_Batch* b = gpu_batch_begin(0, acc_bytes + 8);
      gpu_batch_buffer_zero(b, table, slots_n * 8);
      gpu_batch_buffer_zero(b, flags, 8);
      gpu_batch_buffer_zero(b, acc, acc_bytes);
      if (arg_count && need_minmax) {
        gpu_batch_buffer_fill(b, acc, zero_bytes + R * 4 * arg_count, 0xffffffffu);
        gpu_batch_buffer_zero(b, acc, zero_bytes);
      }
      gpu_batch_kernel_execute(b, kb, (U32)build_threads, QE_WG);
      gpu_batch_kernel_execute(b, kp, blocks * QE_WG, QE_WG);
      gpu_batch_buffer_read(b, flags, fl, 8);
      gpu_batch_buffer_read(b, acc, host_acc, acc_bytes);
      ok = gpu_batch_end(b);
      gpu_hip_kernel_set_lds_bytes(kp, 0);
      if (!ok) qe_set_error("join probe: %s", gpu_hip_last_error());
      if (ok && fl[0]) {
        ok = 0;
        qe_set_error("qe_aggregate_join: duplicate dim keys (the probe form needs unique keys)");
      }
      if (ok && fl[1]) {
        ok = 0;
        qe_set_error("qe_aggregate_join: hash table full");
      }
      if (trace) {
        trace->kernel_count = 2;
        trace->rows_scanned = fact->row_count;
        trace->bytes_read = fact->row_count * 4 * (U64)slots.count;
        snprintf(trace->kernel, sizeof(trace->kernel), "join_probe");
      }
    }
    if (ok) {
      U64* cnt = (U64*)host_acc;
      U64* sums = cnt + R;
      U32* mins = (U32*)(sums + R * arg_count);
      U32* maxs = mins + R * arg_count;
      if (group_mode) {
        for (U64 gi = 0; gi < R; gi++) {
          if (!cnt[gi]) continue;
          memset(gv, 0, sizeof(gv));
          if (K == 1) {
            gv[0] = gmin + (U32)gi;
          } else {
            U64 c = gi;
            for (U32 k = K; k-- > 0;) {
              gv[k] = gk.min_code[k] + (U32)(c % gk.range[k]);
              c /= gk.range[k];
            }
          }
          QE_GroupAcc* g = qe_group_get(&groups, gv);
          qe_group_add(g, funcs, expr_arg, expr_count, arg_type, cnt[gi], sums, mins, maxs, R, gi);
        }
      } else {
        for (U64 v = 0; v < pass_count; v++) {
          U64 d = pass[v];
          if (!cnt[d]) continue;
          QE_DIM_GROUP(d);
          QE_GroupAcc* g = qe_group_get(&groups, gv);
          qe_group_add(g, funcs, expr_arg, expr_count, arg_type, cnt[d], sums, mins, maxs, R, d);
        }
      }
    }
    free(host_acc);
    qe_scratch_release(table);
    qe_scratch_release(flags);
    qe_scratch_release(acc);
    qe_scratch_release(list_buf);
  }
#undef QE_DIM_GROUP
  if (ok) {
    /* ascending group values (codes of I32 columns are biased: signed order) */
    qsort(groups.groups, groups.count, sizeof(QE_GroupAcc), qe_cmp_group);
    for (U64 g = 0; g < groups.count && g < cap; g++) {
      QE_GroupAcc* a = &groups.groups[g];
      if (out_keys)
        for (U32 k = 0; k < K; k++)
          out_keys[g * K + k] = gi32[k] ? (F64)(S32)(a->value[k] ^ 0x80000000u) : (F64)a->value[k];
      for (U32 e = 0; e < expr_count; e++) out_cells[g * expr_count + e] = qe_agg_cell(funcs[e], a->count, a->sum[e], a->mn[e], a->mx[e]);
    }
    *out_group_count = groups.count;
    if (trace) {
      trace->path = path;
      trace->rows_after = groups.count;
    }
    /* the fact-side SUMs reach 2^53: query_exec.c:7028-7049 via :8771 */
    ok = qe_check_sum_exact(fact, funcs, fact_arg_cols, expr_count, out_cells, groups.count < cap ? groups.count : cap);
  }
  qe_group_map_free(&groups);
  qe_scratch_release(pass_dev);
  qe_scratch_release(code_buf);
  free(dkeys);
  for (U32 k = 0; k < QE_AGG_MAX_GROUP_COLS; k++) free(dgroups[k]);
  free(pass);
  qe_trace_end(trace, t0, p0);
  return ok;
}

/* the same join: dim and keys, dim WHERE nodes, group columns, functions,
 * arguments, capacity and whether it asks for its keys */
static B32 qe_join_same(const QE_JoinQuery* a, const QE_JoinQuery* b)
{
  if (a->dim != b->dim || a->fact_key_col != b->fact_key_col || a->dim_key_col != b->dim_key_col ||
      a->dim_where_count != b->dim_where_count || a->group_count != b->group_count ||
      a->expr_count != b->expr_count || a->cap != b->cap || (a->out_keys != 0) != (b->out_keys != 0) ||
      a->group_count > QE_AGG_MAX_GROUP_COLS)
    return 0;
  if (memcmp(a->group_cols, b->group_cols, a->group_count * sizeof(S32))) return 0;
  if (a->dim_where_count &&
      (!a->dim_where || !b->dim_where || memcmp(a->dim_where, b->dim_where, a->dim_where_count * sizeof(QE_IRNode))))
    return 0;
  return memcmp(a->funcs, b->funcs, a->expr_count * 4) == 0 && memcmp(a->arg_cols, b->arg_cols, a->expr_count * 4) == 0;
}

/* run order: joins that can share a fact side next to each other (same dim
 * and keys, then the same argument columns), arrival order otherwise */
static int qe_join_order_cmp(const QE_JoinQuery* queries, U32 ia, U32 ib)
{
  const QE_JoinQuery* a = &queries[ia];
  const QE_JoinQuery* b = &queries[ib];
  if (a->dim != b->dim) return (uintptr_t)a->dim < (uintptr_t)b->dim ? -1 : 1;
  if (a->fact_key_col != b->fact_key_col) return a->fact_key_col < b->fact_key_col ? -1 : 1;
  if (a->dim_key_col != b->dim_key_col) return a->dim_key_col < b->dim_key_col ? -1 : 1;
  U32 sa[QE_AGG_MAX_EXPRS], sb[QE_AGG_MAX_EXPRS], na = 0, nb = 0;
  for (U32 e = 0; e < a->expr_count && e < QE_AGG_MAX_EXPRS; e++)
    if (a->arg_cols[e] >= 0 && a->funcs[e] != QE_AGG_FUNC_COUNT) sa[na++] = (U32)a->arg_cols[e];
  for (U32 e = 0; e < b->expr_count && e < QE_AGG_MAX_EXPRS; e++)
    if (b->arg_cols[e] >= 0 && b->funcs[e] != QE_AGG_FUNC_COUNT) sb[nb++] = (U32)b->arg_cols[e];
  if (na != nb) return na < nb ? -1 : 1;
  for (U32 i = 0; i < na; i++)
    if (sa[i] != sb[i]) return sa[i] < sb[i] ? -1 : 1;
  return ia < ib ? -1 : (ia > ib ? 1 : 0);
}

static void qe_join_copy_result(const QE_JoinQuery* from, QE_JoinQuery* to)
{
  to->ok = from->ok;
  if (!from->ok) return;
  const U64 g = *from->out_group_count;
  *to->out_group_count = g;
  const U64 n = g < to->cap ? g : to->cap;
  if (to->out_keys && from->out_keys) memcpy(to->out_keys, from->out_keys, n * to->group_count * sizeof(F64));
  memcpy(to->out_cells, from->out_cells, n * to->expr_count * sizeof(U64));
}

B32 qe_aggregate_join_batch(QE_Table* fact, QE_JoinQuery* queries, U32 count, QE_JoinBatchTrace* trace)
{
  const U64 t0 = qe_now_ns();
  if (trace) memset(trace, 0, sizeof(*trace));
  if (!fact || (!queries && count)) {
    qe_set_error("qe_aggregate_join_batch: null argument");
    return 0;
  }
  for (U32 i = 0; i < count; i++)
    if (!queries[i].dim || !queries[i].funcs || !queries[i].arg_cols || !queries[i].out_group_count ||
        !queries[i].out_cells) {
      qe_set_error("qe_aggregate_join_batch: query %u has a null argument", i);
      return 0;
    }
  U32* rep_of = (U32*)malloc(sizeof(U32) * (count ? count : 1));
  U32* order = (U32*)malloc(sizeof(U32) * (count ? count : 1));
  if (!rep_of || !order) {
    free(rep_of);
    free(order);
    qe_set_error("qe_aggregate_join_batch: out of memory");
    return 0;
  }
  /* QE_BATCH_DEDUPE (default 1): identical joins run once */
  const B32 dedupe = qe_settings_u64("QE_BATCH_DEDUPE", 1) != 0;
  U32 runs = 0;
  for (U32 i = 0; i < count; i++) {
    rep_of[i] = ~0u;
    for (U32 j = 0; dedupe && j < i; j++)
      if (rep_of[j] == ~0u && qe_join_same(&queries[i], &queries[j])) {
        rep_of[i] = j;
        break;
      }
    if (rep_of[i] == ~0u) order[runs++] = i;
  }
  /* insertion sort (a batch holds at most a server's max_batch joins); the
   * comparator takes the queries as an argument, so concurrent batches on
   * other threads share no state */
  for (U32 i = 1; i < runs; i++) {
    const U32 o = order[i];
    U32 j = i;
    while (j > 0 && qe_join_order_cmp(queries, order[j - 1], o) > 0) {
      order[j] = order[j - 1];
      j--;
    }
    order[j] = o;
  }
  QE_JoinShare share;
  memset(&share, 0, sizeof(share));
  t_join_share = &share;
  char first_error[256] = "";
  for (U32 r = 0; r < runs; r++) {
    QE_JoinQuery* q = &queries[order[r]];
    q->ok = qe_aggregate_join(fact, q->fact_key_col, q->dim, q->dim_key_col, q->dim_where, q->dim_where_count,
                              q->group_cols, q->group_count, q->funcs, q->arg_cols, q->expr_count, q->cap,
                              q->out_group_count, q->out_keys, q->out_cells, 0);
    if (!q->ok && !first_error[0]) snprintf(first_error, sizeof(first_error), "%s", qe_last_error());
  }
  t_join_share = 0;
  const U32 shared = share.fact_passes_shared;
  qe_join_share_drop_part(&share);
  qe_join_share_drop_dense(&share);
  for (U32 i = 0; i < count; i++)
    if (rep_of[i] != ~0u) qe_join_copy_result(&queries[rep_of[i]], &queries[i]);
  free(rep_of);
  free(order);
  if (first_error[0]) qe_set_error("%s", first_error);
  if (trace) {
    trace->queries = count;
    trace->runs = runs;
    trace->fact_passes = runs - shared;
    trace->wall_ns = qe_now_ns() - t0;
  }
  return 1;
}
