/* qe_hash_join: the equi-join that produces row pairs, the general path of
 * ComputeQL's joins (qe_hash_join_impl, query_exec.c:14541-15358, with the
 * kernels hash_join_build_count.comp, csr_scatter.comp and
 * hash_join_probe.comp).
 *
 * What is kept from the reference: keys compare as f64 values whatever the
 * numeric column type (query_exec.c:14704-14725 converts both sides to
 * f64; hash_join_probe.comp compares doubles), every build row is kept so
 * duplicate build keys fan out, a LEFT join pairs an unmatched probe row
 * with all ones (BUILD_NULL, hash_join_probe.comp), SEMI keeps the probe
 * rows with a match and ANTI those without, each once, in probe order
 * (qe_join_filter_left_rows, query_exec.c:14470-14520), and a pair is 16
 * bytes: probe u64, build u64 (the reference's four u32 words).
 *
 * What differs, on purpose:
 * - Pairs come out sorted by probe position, then build position, so the
 *   result is the same on every run (the reference's atomics give any
 *   order; its own tests compare an order-free checksum,
 *   bench_optimizer.c:1846-1890).
 * - -0.0 and 0.0 hash alike (the reference compares them equal but hashes
 *   their bits, so they meet only when their buckets happen to agree).
 * - The build side is a table of distinct key codes with a count per key
 *   (open addressing, 2 slots a build row at least), and each key's build
 *   positions form one contiguous group; a probe row reads one slot and
 *   one group instead of walking a bucket of candidate rows.
 * - The output is sized by a counting pass (per-block totals added here)
 *   instead of a capacity guess and a retry.
 * String keys are not supported: the tables here hold numeric columns (a
 * string column joins through its dictionary codes, a u32 column).
 */
#include "qe_internal.h"

#include <stdlib.h>
#include <string.h>

#define QE_HJ_BLOCKS_PER_CU 4

static B32 qe_hj_type_ok(U32 type)
{
  return type == CQL_ColumnType_U32 || type == CQL_ColumnType_I32 || type == CQL_ColumnType_F32 ||
         type == CQL_ColumnType_U64 || type == CQL_ColumnType_I64 || type == CQL_ColumnType_F64;
}

static U32 qe_hj_grid(U64 items)
{
  U64 g = (items + QE_WG - 1) / QE_WG;
  U64 cap = gpu_hip_compute_units() * QE_HJ_BLOCKS_PER_CU;
  if (g > cap) g = cap;
  return (U32)(g ? g : 1);
}

U32 qe_hash_join(QE_Table* probe, U32 probe_key_col, const U64* probe_rows, U64 probe_count, QE_Table* build,
                 U32 build_key_col, const U64* build_rows, U64 build_count, U32 join_type, U64* out, U64 cap,
                 U64* out_count, QE_Trace* trace)
{
  U64 t0 = 0, p0 = 0;
  qe_trace_begin(trace, &t0, &p0);
  if (!probe || !build || !out_count || (cap && !out)) {
    qe_set_error("qe_hash_join: null argument");
    return 0;
  }
  if (!qe_require_device(probe, "qe_hash_join") || !qe_require_device(build, "qe_hash_join")) return 0;
  if (!qe_require_non_null(probe, (S32)probe_key_col, "qe_hash_join") ||
      !qe_require_non_null(build, (S32)build_key_col, "qe_hash_join"))
    return 0;
  if (probe_key_col >= probe->column_count || build_key_col >= build->column_count) {
    qe_set_error("qe_hash_join: unknown key column");
    return 0;
  }
  const U32 ptype = probe->columns[probe_key_col].type, btype = build->columns[build_key_col].type;
  if (!qe_hj_type_ok(ptype) || !qe_hj_type_ok(btype)) {
    qe_set_error("qe_hash_join: join keys must be numeric columns");
    return 0;
  }
  if (join_type > QE_JoinType_Anti) {
    qe_set_error("qe_hash_join: unknown join type %u", join_type);
    return 0;
  }
  const U64 np = probe_rows ? probe_count : probe->row_count;
  const U64 nb = build_rows ? build_count : build->row_count;
  if (probe_rows)
    for (U64 i = 0; i < np; i++)
      if (probe_rows[i] >= probe->row_count) {
        qe_set_error("qe_hash_join: probe row %llu out of range", (unsigned long long)probe_rows[i]);
        return 0;
      }
  if (build_rows)
    for (U64 i = 0; i < nb; i++)
      if (build_rows[i] >= build->row_count) {
        qe_set_error("qe_hash_join: build row %llu out of range", (unsigned long long)build_rows[i]);
        return 0;
      }
  if (nb >= (1ull << 31)) {
    qe_set_error("qe_hash_join: %llu build rows exceed 2^31", (unsigned long long)nb);
    return 0;
  }
  *out_count = 0;
  /* distinct-key table: T = 2^log2 >= 2 x build rows */
  U32 log2 = 4;
  while ((1ull << log2) < 2 * nb) log2++;
  const U64 T = 1ull << log2;
  const U64 tiles = (T + QE_WG * 8 - 1) / (QE_WG * 8);
  const U32 pgrid = qe_hj_grid(np);

  GPU_Kernel* kb = qe_kernel("hjoin_build");
  GPU_Kernel* ks1 = qe_kernel("scan_u32_blocks");
  GPU_Kernel* ks2 = qe_kernel("scan_u32_top");
  GPU_Kernel* ks3 = qe_kernel("scan_u32_add");
  GPU_Kernel* kc = qe_kernel("hjoin_scatter");
  GPU_Kernel* kss = qe_kernel("hjoin_sort_small");
  GPU_Kernel* ksb = qe_kernel("hjoin_sort_big");
  GPU_Kernel* kp = qe_kernel("hjoin_probe");
  GPU_Buffer* plist = probe_rows && np ? qe_scratch("hj_probe_rows", np * 8) : 0;
  GPU_Buffer* blist = build_rows && nb ? qe_scratch("hj_build_rows", nb * 8) : 0;
  GPU_Buffer* keys = qe_scratch("hj_keys", T * 8);
  GPU_Buffer* cnt = qe_scratch("hj_counts", T * 4);
  GPU_Buffer* start = qe_scratch("hj_starts", T * 4);
  GPU_Buffer* cursor = qe_scratch("hj_cursor", T * 4);
  GPU_Buffer* tilesum = qe_scratch("hj_tiles", tiles * 4);
  GPU_Buffer* slot_of = qe_scratch("hj_slot_of", (nb ? nb : 1) * 4);
  GPU_Buffer* groups = qe_scratch("hj_groups", (nb ? nb : 1) * 4);
  GPU_Buffer* big = qe_scratch("hj_big", (nb ? nb : 1) * 4);
  GPU_Buffer* meta = qe_scratch("hj_meta", 64);
  GPU_Buffer* blk = qe_scratch("hj_block_totals", (U64)pgrid * 8);
  B32 ok = kb && ks1 && ks2 && ks3 && kc && kss && ksb && kp && keys && cnt && start && cursor && tilesum && slot_of &&
           groups && big && meta && blk && (!probe_rows || !np || plist) && (!build_rows || !nb || blist);
  U64* totals = ok ? (U64*)malloc((U64)pgrid * 8) : 0;
  ok = ok && totals;
  U32 kernels = 0;
  if (ok) {
    /* build: distinct keys, counts, scan, groups, sorted groups */
    gpu_kernel_set_arg_buffer(kb, 0, blist);
    if (!blist) gpu_hip_kernel_clear_arg(kb, 0);
    gpu_kernel_set_arg_buffer(kb, 2, keys);
    gpu_kernel_set_arg_buffer(kb, 3, cnt);
    gpu_kernel_set_arg_buffer(kb, 4, slot_of);
    gpu_kernel_set_arg_buffer(kb, 6, build->columns[build_key_col].buffer);
    gpu_kernel_set_arg_u64(kb, 0, nb);
    qe_kernel_set_word(kb, QK_W_KEY, btype);
    qe_kernel_set_word(kb, QK_W_LOG2, log2);
    gpu_kernel_set_arg_buffer(ks1, 0, cnt);
    gpu_kernel_set_arg_buffer(ks1, 1, start);
    gpu_kernel_set_arg_buffer(ks1, 2, tilesum);
    gpu_kernel_set_arg_u64(ks1, 0, T);
    gpu_kernel_set_arg_buffer(ks2, 2, tilesum);
    gpu_kernel_set_arg_u64(ks2, 0, tiles);
    gpu_kernel_set_arg_buffer(ks3, 1, start);
    gpu_kernel_set_arg_buffer(ks3, 2, tilesum);
    gpu_kernel_set_arg_u64(ks3, 0, T);
    gpu_kernel_set_arg_buffer(kc, 4, slot_of);
    gpu_kernel_set_arg_buffer(kc, 5, start);
    gpu_kernel_set_arg_buffer(kc, 7, cursor);
    gpu_kernel_set_arg_buffer(kc, 8, groups);
    gpu_kernel_set_arg_u64(kc, 0, nb);
    gpu_kernel_set_arg_buffer(kss, 3, cnt);
    gpu_kernel_set_arg_buffer(kss, 5, start);
    gpu_kernel_set_arg_buffer(kss, 8, groups);
    gpu_kernel_set_arg_buffer(kss, 9, big);
    gpu_kernel_set_arg_buffer(kss, 10, meta);
    gpu_kernel_set_arg_u64(kss, 0, T);
    gpu_kernel_set_arg_buffer(ksb, 3, cnt);
    gpu_kernel_set_arg_buffer(ksb, 5, start);
    gpu_kernel_set_arg_buffer(ksb, 8, groups);
    gpu_kernel_set_arg_buffer(ksb, 9, big);
    gpu_kernel_set_arg_buffer(ksb, 10, meta);
    /* probe, counting pass */
    gpu_kernel_set_arg_buffer(kp, 0, plist);
    if (!plist) gpu_hip_kernel_clear_arg(kp, 0);
    gpu_kernel_set_arg_buffer(kp, 1, blist);
    if (!blist) gpu_hip_kernel_clear_arg(kp, 1);
    gpu_kernel_set_arg_buffer(kp, 2, keys);
    gpu_kernel_set_arg_buffer(kp, 3, cnt);
    gpu_kernel_set_arg_buffer(kp, 5, start);
    gpu_kernel_set_arg_buffer(kp, 6, probe->columns[probe_key_col].buffer);
    gpu_kernel_set_arg_buffer(kp, 8, groups);
    gpu_hip_kernel_clear_arg(kp, 9);
    gpu_kernel_set_arg_buffer(kp, 10, blk);
    gpu_kernel_set_arg_u64(kp, 0, np);
    gpu_kernel_set_arg_u64(kp, 1, 0);
    gpu_kernel_set_arg_u64(kp, 2, 0);
    qe_kernel_set_word(kp, QK_W_KEY, ptype);
    qe_kernel_set_word(kp, QK_W_LOG2, log2);
    qe_kernel_set_word(kp, QK_W_ARGS, join_type);

    GPU_Batch* b = gpu_batch_begin((plist ? np * 8 : 0) + (blist ? nb * 8 : 0), (U64)pgrid * 8);
    if (plist) gpu_batch_buffer_write(b, plist, (void*)probe_rows, np * 8);
    if (blist) gpu_batch_buffer_write(b, blist, (void*)build_rows, nb * 8);
    gpu_batch_buffer_fill(b, keys, T * 8, 0xffffffffu);
    gpu_batch_buffer_zero(b, cnt, T * 4);
    gpu_batch_buffer_zero(b, cursor, T * 4);
    gpu_batch_buffer_zero(b, meta, 64);
    if (nb) gpu_batch_kernel_execute(b, kb, qe_hj_grid(nb) * QE_WG, QE_WG);
    gpu_batch_kernel_execute(b, ks1, (U32)tiles * QE_WG, QE_WG);
    gpu_batch_kernel_execute(b, ks2, QE_WG, QE_WG);
    gpu_batch_kernel_execute(b, ks3, (U32)tiles * QE_WG, QE_WG);
    if (nb) {
      gpu_batch_kernel_execute(b, kc, qe_hj_grid(nb) * QE_WG, QE_WG);
      gpu_batch_kernel_execute(b, kss, qe_hj_grid(T) * QE_WG, QE_WG);
      gpu_batch_kernel_execute(b, ksb, (U32)(gpu_hip_compute_units() * QE_HJ_BLOCKS_PER_CU) * QE_WG, QE_WG);
    }
    gpu_batch_kernel_execute(b, kp, pgrid * QE_WG, QE_WG);
    gpu_batch_buffer_read(b, blk, totals, (U64)pgrid * 8);
    ok = gpu_batch_end(b);
    kernels = nb ? 8 : 5;
  }
  U64 total = 0;
  B32 too_many = 0;
  if (ok) {
    /* block totals -> each block's first record */
    for (U32 i = 0; i < pgrid; i++) {
      U64 t = totals[i];
      totals[i] = total;
      total += t;
    }
    *out_count = total;
    /* the write pass ranks records with u32 scans */
    if (total >> 32) {
      qe_set_error("qe_hash_join: %llu output records exceed 2^32", (unsigned long long)total);
      too_many = 1;
      ok = 0;
    }
  }
  if (ok) {
    const U64 want = total < cap ? total : cap;
    const U64 rec = join_type >= QE_JoinType_Semi ? 1 : 2;
    if (want) {
      GPU_Buffer* outbuf = qe_scratch("hj_out", want * rec * 8);
      ok = outbuf != 0;
      if (ok) {
        gpu_kernel_set_arg_buffer(kp, 9, outbuf);
        gpu_kernel_set_arg_u64(kp, 1, 1);
        gpu_kernel_set_arg_u64(kp, 2, want);
        GPU_Batch* b = gpu_batch_begin((U64)pgrid * 8, want * rec * 8);
        gpu_batch_buffer_write(b, blk, totals, (U64)pgrid * 8);
        gpu_batch_kernel_execute(b, kp, pgrid * QE_WG, QE_WG);
        gpu_batch_buffer_read(b, outbuf, out, want * rec * 8);
        ok = gpu_batch_end(b);
        gpu_hip_kernel_clear_arg(kp, 9);
        kernels++;
      }
      qe_scratch_release(outbuf);
    }
  }
  free(totals);
  qe_scratch_release(plist);
  qe_scratch_release(blist);
  qe_scratch_release(keys);
  qe_scratch_release(cnt);
  qe_scratch_release(start);
  qe_scratch_release(cursor);
  qe_scratch_release(tilesum);
  qe_scratch_release(slot_of);
  qe_scratch_release(groups);
  qe_scratch_release(big);
  qe_scratch_release(meta);
  qe_scratch_release(blk);
  if (!ok) {
    if (!too_many) qe_set_error("qe_hash_join: GPU dispatch failed: %s", gpu_hip_last_error());
    return 0;
  }
  if (trace) {
    trace->path = QE_AggPath_None;
    trace->kernel_count = kernels;
    trace->rows_before = np;
    trace->rows_scanned = np;
    trace->rows_after = total;
    trace->bytes_read = np * probe->columns[probe_key_col].width + nb * build->columns[build_key_col].width;
    snprintf(trace->kernel, sizeof(trace->kernel), "hjoin_probe");
  }
  qe_trace_end(trace, t0, p0);
  return 1;
}
