/* Row materialisation: column values read through row ids on the device
 * (selection_gather.comp, qe_selection_gather_column and the device
 * selections of query_exec.h:383-415).  qe_gather takes a row list from
 * the caller (at a stride, so the pairs of qe_hash_join serve directly);
 * qe_scan_materialize keeps the ids of a scan on the device and gathers
 * the requested columns of the matching rows in one batch, so no row id
 * crosses PCIe. */
#include "qe_internal.h"

#include <string.h>

static U32 qe_gather_grid(U64 n)
{
  U64 g = (n + QE_WG - 1) / QE_WG;
  U64 cap = gpu_hip_compute_units() * 8;
  if (g > cap) g = cap;
  return (U32)(g ? g : 1);
}

/* gather up to 4 columns per launch from device ids into out[c] */
static B32 qe_gather_run(QE_Table* t, const U32* cols, U32 col_count, GPU_Buffer* ids, U64 n, U64 stride, U64 offset,
                         GPU_Buffer* upload_to, const U64* upload, U64 upload_bytes, void* const* out)
{
  GPU_Kernel* k = qe_kernel("gather");
  if (!k) return 0;
  B32 ok = 1;
  B32 uploaded = 0;
  for (U32 c0 = 0; c0 < col_count && ok; c0 += 4) {
    const U32 m = col_count - c0 < 4 ? col_count - c0 : 4;
    GPU_Buffer* dst[4] = {0};
    U64 down = 0;
    U32 widths = 0;
    for (U32 j = 0; j < m; j++) {
      const QE_Column* col = &t->columns[cols[c0 + j]];
      dst[j] = qe_scratch(j == 0 ? "gather_out0" : (j == 1 ? "gather_out1" : (j == 2 ? "gather_out2" : "gather_out3")),
                          (n ? n : 1) * col->width);
      ok = ok && dst[j];
      widths |= col->width << (8 * j);
      down += n * col->width;
    }
    if (ok) {
      gpu_kernel_set_arg_buffer(k, 0, ids);
      for (U32 j = 0; j < 4; j++) {
        if (j < m) {
          gpu_kernel_set_arg_buffer(k, 6 + j, t->columns[cols[c0 + j]].buffer);
          gpu_kernel_set_arg_buffer(k, 10 + j, dst[j]);
        } else {
          gpu_hip_kernel_clear_arg(k, 6 + j);
          gpu_hip_kernel_clear_arg(k, 10 + j);
        }
      }
      gpu_kernel_set_arg_u64(k, 0, n);
      gpu_kernel_set_arg_u64(k, 1, stride);
      gpu_kernel_set_arg_u64(k, 2, offset);
      qe_kernel_set_word(k, QK_W_ARGS, m);
      qe_kernel_set_word(k, QK_W_TYPES, widths);
      GPU_Batch* b = gpu_batch_begin(uploaded ? 0 : upload_bytes, down);
      if (upload_to && !uploaded) gpu_batch_buffer_write(b, upload_to, (void*)upload, upload_bytes);
      uploaded = 1;
      gpu_batch_kernel_execute(b, k, qe_gather_grid(n) * QE_WG, QE_WG);
      for (U32 j = 0; j < m; j++) gpu_batch_buffer_read(b, dst[j], out[c0 + j], n * t->columns[cols[c0 + j]].width);
      ok = gpu_batch_end(b);
    }
    for (U32 j = 0; j < m; j++) qe_scratch_release(dst[j]);
  }
  if (!ok) qe_set_error("gather: %s", gpu_hip_last_error());
  return ok;
}

static B32 qe_gather_check(QE_Table* t, const U32* cols, U32 col_count, void* const* out)
{
  if (!t || (col_count && (!cols || !out))) {
    qe_set_error("gather: null argument");
    return 0;
  }
  if (!qe_require_device(t, "gather")) return 0;
  for (U32 c = 0; c < col_count; c++)
    if (!qe_require_non_null(t, (S32)cols[c], "gather")) return 0;
  for (U32 c = 0; c < col_count; c++)
    if (cols[c] >= t->column_count || !out[c]) {
      qe_set_error("gather: bad column %u", c);
      return 0;
    }
  return 1;
}

U32 qe_gather(QE_Table* table, const U32* cols, U32 col_count, const U64* rows, U64 count, U64 stride, U64 offset,
              void* const* out)
{
  if (!qe_gather_check(table, cols, col_count, out)) return 0;
  if (stride == 0 || offset >= stride || (count && !rows)) {
    qe_set_error("gather: stride %llu / offset %llu", (unsigned long long)stride, (unsigned long long)offset);
    return 0;
  }
  for (U64 i = 0; i < count; i++) {
    const U64 id = rows[i * stride + offset];
    if (id != ~0ull && id >= table->row_count) {
      qe_set_error("gather: row %llu out of range", (unsigned long long)id);
      return 0;
    }
  }
  if (!count || !col_count) return 1;
  const U64 bytes = ((count - 1) * stride + offset + 1) * 8;
  GPU_Buffer* ids = qe_scratch("gather_rows", bytes);
  B32 ok = ids && qe_gather_run(table, cols, col_count, ids, count, stride, offset, ids, rows, bytes, out);
  qe_scratch_release(ids);
  return ok;
}

U32 qe_scan_materialize(QE_Table* table, const QE_IRNode* where, U32 where_count, const U32* cols, U32 col_count,
                        U64 cap, U64* out_count, void* const* out)
{
  if (!out_count || !qe_gather_check(table, cols, col_count, out)) return 0;
  GPU_Buffer* ids = 0;
  U64 n = 0;
  if (!qe_scan_filter_device(table, where, where_count, &ids, &n)) return 0;
  *out_count = n;
  const U64 want = n < cap ? n : cap;
  B32 ok = 1;
  if (want && col_count) ok = qe_gather_run(table, cols, col_count, ids, want, 1, 0, 0, 0, 0, out);
  qe_scratch_release(ids);
  return ok;
}
