/*
 * qe_core.c - settings, errors, backend access, device-resident tables,
 * zone maps and the synthetic-table generators of the host operator layer.
 *
 * Reference: settings.c:158-242 (settings_u64), cql.h:206-322 (CQL_Column),
 * cql.c:3818 (zone maps), gpu_dispatch.c:70-151 (backend selection).
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdlib.h>
#include <time.h>

#include "qe_internal.h"

/* ------------------------------------------------------------------ */
/* settings                                                             */
/* ------------------------------------------------------------------ */
#define QE_SETTINGS_CAP 128
typedef struct {
  char key[64];
  U64 value;
} QE_Setting;
static QE_Setting g_settings[QE_SETTINGS_CAP];
static U32 g_setting_count;
static pthread_mutex_t g_settings_lock = PTHREAD_MUTEX_INITIALIZER;

void qe_settings_set_u64(const char* key, U64 value)
{
  pthread_mutex_lock(&g_settings_lock);
  for (U32 i = 0; i < g_setting_count; i++) {
    if (strcmp(g_settings[i].key, key) == 0) {
      g_settings[i].value = value;
      pthread_mutex_unlock(&g_settings_lock);
      return;
    }
  }
  if (g_setting_count < QE_SETTINGS_CAP) {
    snprintf(g_settings[g_setting_count].key, sizeof(g_settings[0].key), "%s", key);
    g_settings[g_setting_count].value = value;
    g_setting_count++;
  }
  pthread_mutex_unlock(&g_settings_lock);
}

U64 qe_settings_u64(const char* key, U64 default_value)
{
  U64 v = default_value;
  pthread_mutex_lock(&g_settings_lock);
  for (U32 i = 0; i < g_setting_count; i++) {
    if (strcmp(g_settings[i].key, key) == 0) {
      v = g_settings[i].value;
      break;
    }
  }
  pthread_mutex_unlock(&g_settings_lock);
  return v;
}

void qe_settings_reset(void)
{
  pthread_mutex_lock(&g_settings_lock);
  g_setting_count = 0;
  pthread_mutex_unlock(&g_settings_lock);
}

/* ------------------------------------------------------------------ */
/* errors and time                                                      */
/* ------------------------------------------------------------------ */
static __thread char t_error[512];

void qe_set_error(const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_error, sizeof(t_error), fmt, ap);
  va_end(ap);
}

const char* qe_last_error(void) { return t_error; }

U64 qe_now_ns(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (U64)ts.tv_sec * 1000000000ull + (U64)ts.tv_nsec;
}

void qe_trace_begin(QE_Trace* t, U64* t0, U64* prof0)
{
  *t0 = qe_now_ns();
  *prof0 = gpu_hip_profile_query(0, 0);
  if (t) memset(t, 0, sizeof(*t));
}

void qe_trace_end(QE_Trace* t, U64 t0, U64 prof0)
{
  if (!t) return;
  t->wall_ns = qe_now_ns() - t0;
  U64 p = gpu_hip_profile_query(0, 0);
  t->gpu_kernel_ns = p >= prof0 ? p - prof0 : 0;
}

/* ------------------------------------------------------------------ */
/* backend access                                                       */
/* ------------------------------------------------------------------ */
static GPU_Backend* g_active_backend;

GPU_Backend* qe_gpu(void)
{
  if (!g_active_backend) {
    g_active_backend = gpu_hip_backend();
    g_active_backend->init();
  }
  return g_active_backend;
}

String8 qe_str8(const char* s)
{
  String8 r;
  r.str = (U8*)s;
  r.size = strlen(s);
  return r;
}

GPU_Kernel* qe_kernel(const char* name)
{
  GPU_Kernel* k = gpu_kernel_alloc(qe_str8(name));
  if (!k) qe_set_error("kernel %s unavailable: %s", name, gpu_hip_last_error());
  return k;
}

void qe_kernel_set_word(GPU_Kernel* k, U32 word, U32 value) { gpu_hip_kernel_set_arg_u32(k, word, value); }

GPU_Buffer* qe_scratch(const char* name, U64 bytes)
{
  char key[96];
  snprintf(key, sizeof(key), "%s@%lu", name, (unsigned long)pthread_self());
  return qe_gpu()->buffer_alloc_pooled(qe_str8(key), bytes, GPU_BufferFlag_ReadWrite, 0);
}

void qe_scratch_release(GPU_Buffer* b)
{
  if (b) gpu_buffer_release(b);
}

B32 qe_buffer_read_checked(GPU_Buffer* b, void* out, U64 size)
{
  if (gpu_hip_buffer_read_checked(b, out, size)) return 1;
  qe_set_error("device read failed: %s", gpu_hip_last_error());
  return 0;
}

B32 qe_buffer_write_checked(GPU_Buffer* b, const void* data, U64 size)
{
  if (gpu_hip_buffer_write_checked(b, data, size)) return 1;
  qe_set_error("device write failed: %s", gpu_hip_last_error());
  return 0;
}

B32 qe_check_sum_exact(const QE_Table* t, const U32* funcs, const S32* arg_cols, U32 expr_count, const U64* cells,
                       U64 rows)
{
  if (qe_settings_u64("QE_WIDE_INT_CHECK", 1) == 0) return 1;
  for (U32 e = 0; e < expr_count; e++) {
    if (funcs[e] != QE_AGG_FUNC_SUM || arg_cols[e] < 0 || (U32)arg_cols[e] >= t->column_count) continue;
    const U32 ty = t->columns[arg_cols[e]].type;
    if (ty != CQL_ColumnType_U32 && ty != CQL_ColumnType_I32 && ty != CQL_ColumnType_U64 && ty != CQL_ColumnType_I64)
      continue;
    for (U64 g = 0; g < rows; g++) {
      F64 v;
      memcpy(&v, &cells[g * expr_count + e], 8);
      if (v >= 9007199254740992.0 || v <= -9007199254740992.0) {
        qe_set_error("SUM(%s) reached 2^53, which the engine adds as an inexact double", t->columns[arg_cols[e]].name);
        return 0;
      }
    }
  }
  return 1;
}

U32 qe_setting_blocks_per_cu(void) { return (U32)qe_settings_u64("QE_HIP_BLOCKS_PER_CU", 4); }

U32 qe_stream_blocks(U64 quads)
{
  U64 cus = gpu_hip_compute_units();
  if (cus == 0) cus = 1;
  U64 cap = cus * qe_setting_blocks_per_cu();
  U64 step = (U64)QE_WG * 4; /* quads per block step */
  U64 need = (quads + step - 1) / step;
  if (need < 1) need = 1;
  return (U32)(need < cap ? need : cap);
}

/* ------------------------------------------------------------------ */
/* tables                                                               */
/* ------------------------------------------------------------------ */
U32 qe_type_width(U32 type)
{
  switch (type) {
    case CQL_ColumnType_U32:
    case CQL_ColumnType_I32:
    case CQL_ColumnType_F32: return 4;
    case CQL_ColumnType_U64:
    case CQL_ColumnType_I64:
    case CQL_ColumnType_F64: return 8;
    default: return 0;
  }
}

QE_Table* qe_table_alloc(U64 row_count)
{
  QE_Table* t = (QE_Table*)calloc(1, sizeof(QE_Table));
  if (t) t->row_count = row_count;
  return t;
}

void qe_table_release(QE_Table* table)
{
  if (!table) return;
  for (U32 c = 0; c < table->column_count; c++) {
    QE_Column* col = &table->columns[c];
    if (col->buffer) gpu_buffer_release(col->buffer);
    if (col->enc_buffer) gpu_buffer_release(col->enc_buffer);
    if (col->host_data) {
      qe_resident_drop_column(col->uid);
      gpu_hip_host_free(col->host_data);
    }
    free(col->zone_min);
    free(col->zone_max);
    free(col->super_zone);
  }
  free(table);
}

U64 qe_table_row_count(QE_Table* table) { return table ? table->row_count : 0; }

S32 qe_table_add_column(QE_Table* table, const char* name, U32 type, const void* host_data, const U8* null_flags)
{
  if (!table || table->column_count >= QE_MAX_COLUMNS) {
    qe_set_error("qe_table_add_column: table full");
    return -1;
  }
  if (!qe_refuse_host(table, "qe_table_add_column")) return -1;
  U32 width = qe_type_width(type);
  if (!width) {
    qe_set_error("qe_table_add_column: unsupported type %u", type);
    return -1;
  }
  B32 has_nulls = 0;
  if (null_flags)
    for (U64 i = 0; i < table->row_count && !has_nulls; i++) has_nulls = null_flags[i] != 0;
  QE_Column* col = &table->columns[table->column_count];
  memset(col, 0, sizeof(*col));
  col->has_nulls = has_nulls;
  snprintf(col->name, sizeof(col->name), "%s", name ? name : "");
  col->type = type;
  col->width = width;
  /* 64 bytes of slack past the last row keep every 16-byte quad load of
   * the kernels inside the allocation; a column with NULLs carries its null
   * bitmap after them (16-byte aligned, one bit a row, 16 bytes of slack) */
  U64 bytes = table->row_count * width + 64;
  U64 bitmap_words = (table->row_count + 31) / 32;
  if (has_nulls) {
    col->null_offset = (bytes + 15) & ~15ull;
    bytes = col->null_offset + bitmap_words * 4 + 16;
  }
  col->buffer = gpu_buffer_alloc(bytes, GPU_BufferFlag_ReadWrite | GPU_BufferFlag_DeviceLocal, 0);
  if (!col->buffer) {
    qe_set_error("qe_table_add_column: allocation of %llu bytes failed: %s", (unsigned long long)bytes,
                 gpu_hip_last_error());
    return -1;
  }
  if (host_data && table->row_count && !qe_buffer_write_checked(col->buffer, host_data, table->row_count * width)) {
    gpu_buffer_release(col->buffer);
    memset(col, 0, sizeof(*col));
    return -1;
  }
  if (has_nulls) {
    U32* bits = (U32*)calloc(bitmap_words + 4, 4);
    GPU_Buffer* view = bits ? gpu_hip_buffer_view(col->buffer, col->null_offset, bitmap_words * 4 + 16) : 0;
    if (!view) {
      free(bits);
      gpu_buffer_release(col->buffer);
      memset(col, 0, sizeof(*col));
      qe_set_error("qe_table_add_column: null bitmap: out of memory");
      return -1;
    }
    for (U64 i = 0; i < table->row_count; i++)
      if (null_flags[i]) bits[i >> 5] |= 1u << (i & 31);
    const B32 wrote = qe_buffer_write_checked(view, bits, bitmap_words * 4 + 16);
    gpu_buffer_release(view);
    free(bits);
    if (!wrote) {
      gpu_buffer_release(col->buffer);
      memset(col, 0, sizeof(*col));
      return -1;
    }
  }
  S32 index = (S32)table->column_count++;
  if (host_data && !qe_table_build_zonemap(table, (U32)index)) return -1;
  return index;
}

/* The operators here do not evaluate NULLs (the reference's GPU kernels do
 * not either: its bytecode folds IS NULL to a constant, query_exec.c:
 * 234-248, and its aggregates refuse nullable columns, :10312-10333).  A
 * query that touches a column with NULLs is refused rather than answered
 * on the stored values; other columns of the table stay usable. */
B32 qe_require_non_null(const QE_Table* t, S32 column, const char* op)
{
  if (t && column >= 0 && (U32)column < t->column_count && t->columns[column].has_nulls) {
    qe_set_error("%s: column '%s' has NULLs; NULL-aware evaluation is not supported on this path", op,
                 t->columns[column].name);
    return 0;
  }
  return 1;
}

B32 qe_where_non_null(const QE_Table* t, const QE_IRNode* where, U32 count, const char* op)
{
  /* a WHERE the leaf kernels evaluate (AND / OR of column-literal
   * comparisons, IN lists, IS [NOT] NULL, at most 8 leaves) evaluates NULLs
   * as the reference does (a NULL row fails its leaf, query_exec.c:
   * 13873-14043); other WHERE shapes over a column with NULLs are refused */
  B32 nullable = 0;
  for (U32 i = 0; where && i < count; i++)
    if (where[i].type == QE_IR_COLUMN && where[i].column < t->column_count && t->columns[where[i].column].has_nulls)
      nullable = 1;
  if (!nullable) return 1;
  static __thread QE_Where w;
  if (qe_where_build(t, where, count, &w)) return 1;
  qe_set_error("%s: a WHERE over a column with NULLs must be an AND / OR of at most %d column-literal comparisons "
               "(%d conjunctions) on this path", op, QE_SCAN_MAX_LEAVES, QE_WHERE_MAX_TERMS);
  return 0;
}

void qe_leaf_nulls(U32* words, const QE_Table* t, U32 column)
{
  const QE_Column* c = &t->columns[column];
  words[5] = c->has_nulls ? (U32)(c->null_offset / 16) : 0u;
}

GPU_Buffer* qe_table_column_buffer(QE_Table* table, U32 column)
{
  if (!table || column >= table->column_count) return 0;
  return table->columns[column].buffer;
}

B32 qe_table_column_read(QE_Table* table, U32 column, U64 row0, U64 rows, void* out)
{
  if (!table || column >= table->column_count || row0 + rows > table->row_count) {
    qe_set_error("qe_table_column_read: bad range");
    return 0;
  }
  QE_Column* col = &table->columns[column];
  if (col->host_data) {
    memcpy(out, (const U8*)col->host_data + row0 * col->width, rows * col->width);
    return 1;
  }
  if (row0 == 0) return qe_buffer_read_checked(col->buffer, out, rows * col->width);
  /* the rows [row0, row0 + rows) as a view of the column */
  GPU_Buffer* view = gpu_hip_buffer_view(col->buffer, row0 * col->width, rows * col->width);
  if (!view) {
    qe_set_error("qe_table_column_read: bad range");
    return 0;
  }
  const B32 ok = qe_buffer_read_checked(view, out, rows * col->width);
  gpu_buffer_release(view);
  return ok;
}

B32 qe_table_build_zonemap(QE_Table* table, U32 column)
{
  if (!table || column >= table->column_count || !qe_refuse_host(table, "qe_table_build_zonemap")) return 0;
  QE_Column* col = &table->columns[column];
  /* the data changed: an encoded copy is stale */
  if (col->enc_buffer) gpu_buffer_release(col->enc_buffer);
  col->enc_buffer = 0;
  col->enc = 0;
  col->enc_base = 0;
  U64 zones = (table->row_count + (1ull << QE_ZONE_SHIFT) - 1) >> QE_ZONE_SHIFT;
  free(col->zone_min);
  free(col->zone_max);
  col->zone_min = col->zone_max = 0;
  free(col->super_zone);
  col->super_zone = 0;
  col->super_count = 0;
  col->zone_count = zones;
  col->zonemap_valid = 0;
  col->min_value = 0.0;
  col->max_value = 0.0;
  if (zones == 0) {
    col->zonemap_valid = 1;
    return 1;
  }
  col->zone_min = (F64*)malloc(sizeof(F64) * zones);
  col->zone_max = (F64*)malloc(sizeof(F64) * zones);
  GPU_Buffer* bmin = qe_scratch("zonemap_min", zones * 8);
  GPU_Buffer* bmax = qe_scratch("zonemap_max", zones * 8);
  GPU_Kernel* k = qe_kernel("zonemap_build");
  if (!bmin || !bmax || !k || !col->zone_min || !col->zone_max) {
    qe_scratch_release(bmin);
    qe_scratch_release(bmax);
    qe_set_error("qe_table_build_zonemap: setup failed");
    return 0;
  }
  gpu_kernel_set_arg_buffer(k, 6, col->buffer);
  gpu_kernel_set_arg_buffer(k, 1, bmin);
  gpu_kernel_set_arg_buffer(k, 2, bmax);
  gpu_kernel_set_arg_u64(k, 0, table->row_count);
  qe_kernel_set_word(k, QK_W_ZTYPE, col->type);
  qe_kernel_set_word(k, QK_W_ZSHIFT, QE_ZONE_SHIFT);
  U64 blocks = zones < 65536 ? zones : 65536;
  GPU_Batch* b = gpu_batch_begin(0, zones * 16);
  gpu_batch_kernel_execute(b, k, (U32)(blocks * QE_WG), QE_WG);
  gpu_batch_buffer_read(b, bmin, col->zone_min, zones * 8);
  gpu_batch_buffer_read(b, bmax, col->zone_max, zones * 8);
  B32 ok = gpu_batch_end(b);
  qe_scratch_release(bmin);
  qe_scratch_release(bmax);
  if (!ok) {
    qe_set_error("qe_table_build_zonemap: %s", gpu_hip_last_error());
    return 0;
  }
  F64 mn = col->zone_min[0], mx = col->zone_max[0];
  for (U64 z = 1; z < zones; z++) {
    if (col->zone_min[z] < mn) mn = col->zone_min[z];
    if (col->zone_max[z] > mx) mx = col->zone_max[z];
  }
  col->min_value = mn;
  col->max_value = mx;
  col->zonemap_valid = 1;
  qe_zonemap_build_super(col);
  return 1;
}

/* ------------------------------------------------------------------ */
/* narrow encoding (qe_column_encoding, query_exec.c:1555-1589)         */
/* ------------------------------------------------------------------ */
U32 qe_col_enc(const QE_Table* t, U32 col)
{
  return qe_settings_u64("QE_ENCODE", 1) ? t->columns[col].enc : 0;
}

GPU_Buffer* qe_col_kbuf(const QE_Table* t, U32 col)
{
  const QE_Column* c = &t->columns[col];
  U32 e = qe_col_enc(t, col);
  return (e == 1 || e == 2) ? c->enc_buffer : c->buffer;
}

void qe_leaf_encode(U32* words, U32 enc, U32 base)
{
  if (!enc || !(words[1] & QE_SCAN_INT_FLAG)) return;
  U32 lo = words[3] - base;
  if (words[1] & QE_SCAN_INT_BIAS) {
    lo -= 0x80000000u;
    words[1] &= ~QE_SCAN_INT_BIAS;
  }
  words[3] = lo;
  words[1] |= enc << 8;
}

/* one u32 flag word set by a kernel, read back */
static B32 qe_flag_kernel(GPU_Kernel* k, QE_Column* col, U64 rows, U32* flag)
{
  GPU_Buffer* f = qe_scratch("enc_flag", 64);
  if (!f) return 0;
  gpu_kernel_set_arg_buffer(k, 6, col->buffer);
  gpu_kernel_set_arg_buffer(k, 3, f);
  gpu_kernel_set_arg_u64(k, 0, rows);
  U64 blocks = (rows / 4 + QE_WG - 1) / QE_WG;
  U64 cap = (U64)gpu_hip_compute_units() * 8;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  GPU_Batch* b = gpu_batch_begin(0, 4);
  gpu_batch_buffer_zero(b, f, 64);
  gpu_batch_kernel_execute(b, k, (U32)(blocks * QE_WG), QE_WG);
  gpu_batch_buffer_read(b, f, flag, 4);
  B32 ok = gpu_batch_end(b);
  qe_scratch_release(f);
  if (!ok) qe_set_error("encode: %s", gpu_hip_last_error());
  return ok;
}

S32 qe_table_encode_column(QE_Table* table, U32 column, B32 allow_sequence)
{
  if (!table || column >= table->column_count) {
    qe_set_error("qe_table_encode_column: bad column");
    return -1;
  }
  if (!qe_refuse_host(table, "qe_table_encode_column") || !qe_require_non_null(table, (S32)column, "qe_table_encode_column"))
    return -1;
  QE_Column* col = &table->columns[column];
  const U64 rows = table->row_count;
  if (col->enc) return (S32)col->enc;
  if (rows == 0 || (col->type != CQL_ColumnType_U32 && col->type != CQL_ColumnType_I32)) return 0;
  /* a sequence: value = first value + row, the last one in the type's range */
  if (allow_sequence && qe_settings_u64("QE_ENCODE_SEQUENCE", 1) && rows < (1ull << 32)) {
    U32 first = 0;
    if (!qe_buffer_read_checked(col->buffer, &first, 4)) return -1;
    B32 fits = col->type == CQL_ColumnType_I32 ? (S64)(S32)first + (S64)rows - 1 <= 2147483647ll
                                               : (U64)first + rows - 1 <= 0xffffffffull;
    GPU_Kernel* k = qe_kernel("seq_check");
    U32 bad = 1;
    if (fits && k) {
      qe_kernel_set_word(k, QK_W_KEY_MIN, first);
      if (!qe_flag_kernel(k, col, rows, &bad)) return -1;
    }
    if (fits && !bad) {
      col->enc = 3;
      col->enc_base = first;
      return 3;
    }
  }
  if (!col->zonemap_valid) return 0;
  const F64 lo = col->min_value, hi = col->max_value;
  const U32 width = hi - lo < 256.0 ? 1 : (hi - lo < 65536.0 ? 2 : 0);
  if (!width) return 0;
  const U32 base = (U32)(S64)lo;
  GPU_Kernel* k = qe_kernel("encode_column");
  /* codes of 4 rows a quad, with 64 bytes of slack past the last quad */
  U64 bytes = ((rows + 3) / 4) * 4 * width + 64;
  GPU_Buffer* out = gpu_buffer_alloc(bytes, GPU_BufferFlag_ReadWrite | GPU_BufferFlag_DeviceLocal, 0);
  if (!k || !out) {
    if (out) gpu_buffer_release(out);
    qe_set_error("qe_table_encode_column: allocation failed");
    return -1;
  }
  gpu_kernel_set_arg_buffer(k, 2, out);
  qe_kernel_set_word(k, QK_W_ARGS, width);
  qe_kernel_set_word(k, QK_W_KEY_MIN, base);
  U32 bad = 1;
  if (!qe_flag_kernel(k, col, rows, &bad)) {
    gpu_buffer_release(out);
    return -1;
  }
  if (bad) {
    gpu_buffer_release(out);
    qe_set_error("qe_table_encode_column: values outside the zone-map range (stale zone map)");
    return -1;
  }
  col->enc_buffer = out;
  col->enc = width;
  col->enc_base = base;
  return (S32)width;
}

/* ------------------------------------------------------------------ */
/* generators                                                           */
/* ------------------------------------------------------------------ */
static B32 qe_run_gen(QE_Table* table, GPU_Kernel* k, U64 rows)
{
  U64 blocks = (rows + QE_WG - 1) / QE_WG;
  U64 cap = (U64)gpu_hip_compute_units() * 16;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  GPU_Batch* b = gpu_batch_begin(0, 0);
  gpu_batch_kernel_execute(b, k, (U32)(blocks * QE_WG), QE_WG);
  if (!gpu_batch_end(b)) {
    qe_set_error("generator failed: %s", gpu_hip_last_error());
    return 0;
  }
  (void)table;
  return 1;
}

B32 qe_gen_events(QE_Table* table, U64 seed, S32 id_col, S32 region_col, S32 amount_col, S32 ts_col)
{
  if (!table || !qe_refuse_host(table, "qe_gen")) return 0;
  GPU_Kernel* k = qe_kernel("gen_events");
  if (!k) return 0;
  S32 cols[4] = {id_col, region_col, amount_col, ts_col};
  for (U32 i = 0; i < 4; i++) {
    if (cols[i] >= 0 && ((U32)cols[i] >= table->column_count || table->columns[cols[i]].type != CQL_ColumnType_U32)) {
      qe_set_error("qe_gen_events: column %d must be a U32 column", cols[i]);
      return 0;
    }
  }
  for (U32 i = 0; i < 4; i++) {
    if (cols[i] >= 0) gpu_kernel_set_arg_buffer(k, 6 + i, table->columns[cols[i]].buffer);
    else gpu_hip_kernel_clear_arg(k, 6 + i); /* the kernel skips absent outputs */
  }
  gpu_kernel_set_arg_u64(k, 0, table->row_count);
  gpu_kernel_set_arg_u64(k, 1, seed);
  gpu_kernel_set_arg_u64(k, 2, 0);
  if (!qe_run_gen(table, k, table->row_count)) return 0;
  for (U32 i = 0; i < 4; i++)
    if (cols[i] >= 0 && !qe_table_build_zonemap(table, (U32)cols[i])) return 0;
  return 1;
}

B32 qe_gen_scale_events(QE_Table* table, S32 id_col, S32 region_col, S32 amount_col, S32 key_col)
{
  if (!table || !qe_refuse_host(table, "qe_gen")) return 0;
  GPU_Kernel* k = qe_kernel("gen_scale_events");
  if (!k) return 0;
  S32 cols[4] = {id_col, region_col, amount_col, key_col};
  for (U32 i = 0; i < 4; i++) {
    if (cols[i] >= 0 && ((U32)cols[i] >= table->column_count || table->columns[cols[i]].type != CQL_ColumnType_U32)) {
      qe_set_error("qe_gen_scale_events: column %d must be a U32 column", cols[i]);
      return 0;
    }
  }
  for (U32 i = 0; i < 4; i++) {
    if (cols[i] >= 0) gpu_kernel_set_arg_buffer(k, 6 + i, table->columns[cols[i]].buffer);
    else gpu_hip_kernel_clear_arg(k, 6 + i);
  }
  gpu_kernel_set_arg_u64(k, 0, table->row_count);
  gpu_kernel_set_arg_u64(k, 2, 0);
  if (!qe_run_gen(table, k, table->row_count)) return 0;
  for (U32 i = 0; i < 4; i++)
    if (cols[i] >= 0 && !qe_table_build_zonemap(table, (U32)cols[i])) return 0;
  return 1;
}

/* dimension rows: keys are a Fisher-Yates permutation of 0..n-1 (draws of
 * the splitmix64 stream seeded with seed ^ 0x2545F4914F6CDD1D), optionally
 * spread over the u32 range by the bijection k -> k * 0x85EBCA77 +
 * 0x165667B1; attr = d(j) % attr_range with seed ^ 0x14057B7EF767814F.
 * Built on the host (n is small) and uploaded. */
static U64 qe_splitmix64_at(U64 seed, U64 k)
{
  U64 z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

B32 qe_gen_dim(QE_Table* table, U64 seed, U32 attr_range, B32 spread_keys, S32 key_col, S32 attr_col)
{
  if (!table || key_col < 0 || (U32)key_col >= table->column_count || table->columns[key_col].type != CQL_ColumnType_U32 ||
      (attr_col >= 0 && ((U32)attr_col >= table->column_count || table->columns[attr_col].type != CQL_ColumnType_U32)) ||
      (attr_col >= 0 && attr_range == 0)) {
    qe_set_error("qe_gen_dim: bad arguments");
    return 0;
  }
  if (!qe_refuse_host(table, "qe_gen_dim")) return 0;
  U64 n = table->row_count;
  U32* keys = (U32*)malloc(sizeof(U32) * (n ? n : 1));
  if (!keys) return 0;
  for (U64 i = 0; i < n; i++) keys[i] = (U32)i;
  U64 ps = seed ^ 0x2545F4914F6CDD1DULL;
  for (U64 i = n; i > 1; i--) {
    U64 j = qe_splitmix64_at(ps, i - 1) % i;
    U32 tmp = keys[i - 1];
    keys[i - 1] = keys[j];
    keys[j] = tmp;
  }
  if (spread_keys)
    for (U64 i = 0; i < n; i++) keys[i] = keys[i] * 0x85EBCA77u + 0x165667B1u;
  B32 wrote = qe_buffer_write_checked(table->columns[key_col].buffer, keys, n * 4);
  if (wrote && attr_col >= 0) {
    U64 as = seed ^ 0x14057B7EF767814FULL;
    for (U64 i = 0; i < n; i++) keys[i] = (U32)(qe_splitmix64_at(as, i) % attr_range);
    wrote = qe_buffer_write_checked(table->columns[attr_col].buffer, keys, n * 4);
  }
  free(keys);
  if (!wrote) return 0;
  if (!qe_table_build_zonemap(table, (U32)key_col)) return 0;
  return attr_col < 0 || qe_table_build_zonemap(table, (U32)attr_col);
}

B32 qe_gen_fact_fk(QE_Table* table, U32 fk_col, U64 seed, QE_Table* dim, U32 dim_key_col)
{
  if (!table || !dim || fk_col >= table->column_count || dim_key_col >= dim->column_count ||
      table->columns[fk_col].type != CQL_ColumnType_U32 || dim->columns[dim_key_col].type != CQL_ColumnType_U32 ||
      dim->row_count == 0) {
    qe_set_error("qe_gen_fact_fk: bad arguments");
    return 0;
  }
  if (!qe_refuse_host(table, "qe_gen_fact_fk") || !qe_require_device(dim, "qe_gen_fact_fk")) return 0;
  GPU_Kernel* k = qe_kernel("gen_fact_fk");
  if (!k) return 0;
  gpu_kernel_set_arg_buffer(k, 6, table->columns[fk_col].buffer);
  gpu_kernel_set_arg_buffer(k, 0, dim->columns[dim_key_col].buffer);
  gpu_kernel_set_arg_u64(k, 0, table->row_count);
  gpu_kernel_set_arg_u64(k, 1, seed);
  gpu_kernel_set_arg_u64(k, 2, 0);
  gpu_kernel_set_arg_u64(k, 3, dim->row_count);
  if (!qe_run_gen(table, k, table->row_count)) return 0;
  return qe_table_build_zonemap(table, fk_col);
}

/* ------------------------------------------------------------------ */
/* result typing: COUNT -> U64, SUM/AVG/MIN/MAX -> F64                  */
/* (qe_agg_func_policy, query_exec.h:917-926; qe_aggregate_finish,      */
/* query_exec.c:7269-7293)                                              */
/* ------------------------------------------------------------------ */
__thread B32 qe_t_raw_int_sums;

/* ------------------------------------------------------------------ */
/* bandwidth probe                                                      */
/* ------------------------------------------------------------------ */
B32 qe_bandwidth_copy(GPU_Buffer* src, GPU_Buffer* dst, U64 bytes, U32 blocks, QE_Trace* trace)
{
  U64 t0, p0;
  qe_trace_begin(trace, &t0, &p0);
  if (!src || !dst || bytes > gpu_hip_buffer_size(src) || bytes > gpu_hip_buffer_size(dst)) {
    qe_set_error("bandwidth_copy: %llu bytes do not fit both buffers", (unsigned long long)bytes);
    return 0;
  }
  GPU_Kernel* k = qe_kernel("bandwidth_copy");
  if (!k) return 0;
  U64 units = bytes / 16;
  if (!blocks) blocks = qe_stream_blocks(units);
  gpu_kernel_set_arg_buffer(k, 0, src);
  gpu_kernel_set_arg_buffer(k, 1, dst);
  gpu_kernel_set_arg_u64(k, 0, units);
  GPU_Batch* b = gpu_batch_begin(0, 0);
  gpu_batch_kernel_execute(b, k, blocks * QE_WG, QE_WG);
  B32 ok = gpu_batch_end(b);
  if (trace) {
    trace->kernel_count = 1;
    trace->bytes_read = units * 16;
    snprintf(trace->kernel, sizeof(trace->kernel), "bandwidth_copy");
  }
  qe_trace_end(trace, t0, p0);
  if (!ok) qe_set_error("bandwidth_copy: %s", gpu_hip_last_error());
  return ok;
}

B32 qe_bandwidth_read(GPU_Buffer* buffer, U64 bytes, U32 blocks, U64* out_sink, QE_Trace* trace)
{
  U64 t0, p0;
  qe_trace_begin(trace, &t0, &p0);
  GPU_Kernel* k = qe_kernel("bandwidth_read");
  GPU_Buffer* sink = qe_scratch("bw_sink", 64);
  if (!k || !sink) return 0;
  U64 units = bytes / 16;
  if (!blocks) blocks = qe_stream_blocks(units);
  gpu_kernel_set_arg_buffer(k, 0, buffer);
  gpu_kernel_set_arg_buffer(k, 1, sink);
  gpu_kernel_set_arg_u64(k, 0, units);
  U64 s = 0;
  GPU_Batch* b = gpu_batch_begin(0, 8);
  gpu_batch_kernel_execute(b, k, blocks * QE_WG, QE_WG);
  gpu_batch_buffer_read(b, sink, &s, 8);
  B32 ok = gpu_batch_end(b);
  qe_scratch_release(sink);
  if (out_sink) *out_sink = s;
  if (trace) {
    trace->kernel_count = 1;
    trace->bytes_read = units * 16;
    snprintf(trace->kernel, sizeof(trace->kernel), "bandwidth_read");
  }
  qe_trace_end(trace, t0, p0);
  if (!ok) qe_set_error("bandwidth_read: %s", gpu_hip_last_error());
  return ok;
}
