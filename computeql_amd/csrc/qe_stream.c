/*
 * qe_stream.c - host-backed tables, the resident chunk cache and streamed
 * execution (the reference's cold path, query_exec.c:1674-2491).
 *
 * A host-backed table keeps its rows in pinned host memory.  Operators over
 * it run chunk by chunk: a chunk is QE_SCAN_STREAM_CHUNK_ROWS rows (a multiple
 * of the 8192-row zone), and each column a query reads is brought into HBM
 * per chunk through the resident chunk cache:
 *
 *   cache          entries keyed by (column uid, chunk index); a budget of
 *                  QE_RESIDENT_BUDGET_PERCENT of HBM (70, capped by
 *                  QE_RESIDENT_BUDGET_BYTES and by what the device has
 *                  free, less 512 MB); victims: chunks used once, most
 *                  recent first, then the least recently used of the rest
 *                  (qe_resident_pick_victim, query_exec.c:1765-1790);
 *                  entries are pinned while a query reads them and become
 *                  visible only once their copy has completed
 *                  (qe_resident_acquire/unpin, query_exec.c:1840-1968)
 *   pipeline       chunk i+1's copies run on a copy stream (DMA from the
 *                  pinned rows) while chunk i's kernels run on the compute
 *                  stream (the stream slots of qe_scan_filter_streamed,
 *                  query_exec.c:2023-2491); a chunk whose zones are all
 *                  pruned is neither copied nor run
 *   operators      each chunk is a table view (row count, zone maps and
 *                  buffers of the chunk) passed to the same scan and
 *                  aggregate drivers as a device-resident table, so the
 *                  kernels and their path choice are the resident ones
 *
 * Results are those of the same query over the table held whole in HBM:
 * row ids ascending; 32-bit integer sums merged as exact 64-bit integers
 * (each chunk's SUM cells carry the raw integer sum); MIN/MAX and COUNT
 * exact; F32/F64/U64/I64 sums merged in F64 in chunk order (the resident
 * path orders the same additions by block instead, so those sums agree to
 * rounding).  Grouped results come out in ascending key order.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdlib.h>

#include "qe_internal.h"

#define QE_RESIDENT_MAX_ENTRIES 4096

#define QE_WHOLE_COLUMN (~0ull)

typedef struct QE_ResidentEntry {
  B32 used;
  B32 ready; /* a chunk's copy has completed; for a whole column, every chunk's */
  U64 uid;
  U64 chunk; /* chunk index, or QE_WHOLE_COLUMN: one buffer for all the rows */
  U64 chunk_rows; /* the chunk size the index and the chunk states refer to */
  GPU_Buffer* buffer;
  U64 bytes;
  U64 last_use;
  U32 uses;
  U32 pins;
  /* whole-column entries: per chunk 0 empty, 1 being copied, 2 filled */
  U64 nchunks, nfilled;
  U8* state;
} QE_ResidentEntry;

static struct {
  QE_ResidentEntry entries[QE_RESIDENT_MAX_ENTRIES];
  U64 tick, bytes_in_use, entry_count;
  U64 hit_bytes, miss_bytes, evicted_bytes;
} g_res;
static pthread_mutex_t g_res_lock = PTHREAD_MUTEX_INITIALIZER;
static __thread B32 qe_t_res_locked;
static U64 g_next_uid = 1;

static void res_lock(void)
{
  pthread_mutex_lock(&g_res_lock);
  qe_t_res_locked = 1;
}

static void res_unlock(void)
{
  qe_t_res_locked = 0;
  pthread_mutex_unlock(&g_res_lock);
}

/* rows of a streamed chunk: QE_SCAN_STREAM_CHUNK_ROWS (16,777,216,
 * query_exec.c:2096-2099; QE_STREAM_CHUNK_ROWS is accepted as an older
 * name), at most GPU_MAX_BUFFER_SIZE bytes of the widest column read when
 * that is set (query_exec.c:2740-2756 sizes chunks by it; here columns are
 * otherwise whole in HBM, so it has no default), a multiple of the
 * 8192-row zone so a chunk's zone maps are a slice of the column's */
static U64 qe_stream_chunk_rows(U32 widest)
{
  U64 zone = 1ull << QE_ZONE_SHIFT;
  U64 r = qe_settings_u64("QE_SCAN_STREAM_CHUNK_ROWS", qe_settings_u64("QE_STREAM_CHUNK_ROWS", 16777216ull));
  U64 max_bytes = qe_settings_u64("GPU_MAX_BUFFER_SIZE", 0);
  if (max_bytes && widest && max_bytes / widest < r) r = max_bytes / widest;
  if (r < zone) return zone;
  return r & ~(zone - 1);
}

/* chunk slots in flight: QE_SCAN_STREAM_SLOTS (3, clamped to 2..8,
 * query_exec.c:2253): the copies of the next slots - 1 chunks run while a
 * chunk's kernels do */
#define QE_STREAM_MAX_SLOTS 8
static U32 qe_stream_slots(void)
{
  U64 n = qe_settings_u64("QE_SCAN_STREAM_SLOTS", 3);
  return (U32)(n < 2 ? 2 : (n > QE_STREAM_MAX_SLOTS ? QE_STREAM_MAX_SLOTS : n));
}

/* callers hold the lock */
static U64 res_budget_locked(void)
{
  U64 percent = qe_settings_u64("QE_RESIDENT_BUDGET_PERCENT", 70);
  if (percent > 100) percent = 100;
  U64 budget = gpu_hip_device_total_memory() / 100 * percent;
  U64 cap = qe_settings_u64("QE_RESIDENT_BUDGET_BYTES", 0);
  if (cap && cap < budget) budget = cap;
  /* other allocations, the runtime and other processes share the device */
  U64 reserve = 512ull << 20;
  U64 reachable = g_res.bytes_in_use + gpu_hip_device_free_memory();
  reachable = reachable > reserve ? reachable - reserve : 0;
  return budget < reachable ? budget : reachable;
}

static void res_free_entry(QE_ResidentEntry* e)
{
  gpu_buffer_release(e->buffer);
  free(e->state);
  g_res.bytes_in_use -= e->bytes < g_res.bytes_in_use ? e->bytes : g_res.bytes_in_use;
  g_res.entry_count--;
  memset(e, 0, sizeof(*e));
}

static QE_ResidentEntry* res_pick_victim(void)
{
  QE_ResidentEntry* probation = 0; /* used once: most recent first */
  QE_ResidentEntry* lru = 0;       /* used again: least recent first */
  for (U32 i = 0; i < QE_RESIDENT_MAX_ENTRIES; i++) {
    QE_ResidentEntry* e = &g_res.entries[i];
    if (!e->used || e->pins) continue;
    if (e->uses <= 1) {
      if (!probation || e->last_use > probation->last_use) probation = e;
    } else if (!lru || e->last_use < lru->last_use) {
      lru = e;
    }
  }
  return probation ? probation : lru;
}

static U64 res_evict_locked(U64 wanted)
{
  U64 freed = 0;
  while (freed < wanted) {
    QE_ResidentEntry* v = res_pick_victim();
    if (!v) break;
    freed += v->bytes;
    g_res.evicted_bytes += v->bytes;
    res_free_entry(v);
  }
  return freed;
}

/* the allocator's last resort when HBM runs out; a thread already inside
 * the cache (allocating a chunk) gets nothing, its caller evicted first */
static U64 res_evict_hook(U64 wanted)
{
  if (qe_t_res_locked) return 0;
  res_lock();
  U64 freed = res_evict_locked(wanted);
  res_unlock();
  return freed;
}

/* room for `bytes` under the budget, evicting what is not pinned */
static B32 res_room_locked(U64 bytes)
{
  U64 budget = res_budget_locked();
  if (bytes > budget) return 0;
  if (g_res.bytes_in_use + bytes <= budget) return 1;
  U64 over = g_res.bytes_in_use + bytes - budget;
  return res_evict_locked(over) >= over;
}

static QE_ResidentEntry* res_new_entry_locked(QE_ResidentEntry* slot, U64 uid, U64 chunk, U64 chunk_rows, U64 bytes)
{
  GPU_Buffer* buf = gpu_buffer_alloc(bytes, GPU_BufferFlag_ReadWrite | GPU_BufferFlag_DeviceLocal, 0);
  if (!buf) return 0;
  memset(slot, 0, sizeof(*slot));
  slot->used = 1;
  slot->uid = uid;
  slot->chunk = chunk;
  slot->chunk_rows = chunk_rows;
  slot->buffer = buf;
  slot->bytes = bytes;
  slot->last_use = ++g_res.tick;
  slot->uses = 1;
  g_res.bytes_in_use += bytes;
  g_res.entry_count++;
  return slot;
}

/* The buffer chunk `chunk` of column `col` is read from, pinned in the
 * cache, and whether its rows must still be copied in; 0 when the cache
 * cannot hold it (the caller then streams through a buffer of its own).
 * While the whole column fits the budget (QE_RESIDENT_WHOLE, default 1)
 * its chunks share one buffer of the column's size, which a warm query
 * reads in one pass; otherwise each chunk is an entry of its own.  A chunk
 * of a whole-column entry comes back as a view of that buffer (*view, the
 * caller's to release).  Every entry returned goes back through
 * res_unpin. */
static QE_ResidentEntry* res_acquire(const QE_Column* col, U64 rows, U64 chunk, U64 chunk_rows, B32* needs_copy,
                                     GPU_Buffer** view)
{
  *needs_copy = 0;
  *view = 0;
  const U64 c0 = chunk * chunk_rows;
  const U64 n = rows - c0 < chunk_rows ? rows - c0 : chunk_rows;
  const U64 bytes = n * col->width + 64; /* slack for the kernels' 16-byte loads */
  if (qe_settings_u64("QE_RESIDENT", 1) == 0 || n == 0) return 0;
  gpu_hip_set_evict_hook(res_evict_hook);
  res_lock();
  QE_ResidentEntry *free_slot = 0, *whole = 0, *got = 0;
  for (U32 i = 0; i < QE_RESIDENT_MAX_ENTRIES; i++) {
    QE_ResidentEntry* e = &g_res.entries[i];
    if (!e->used) {
      if (!free_slot) free_slot = e;
      continue;
    }
    if (e->uid != col->uid) continue;
    if (e->chunk == QE_WHOLE_COLUMN) whole = e;
    else if (e->chunk == chunk && e->chunk_rows == chunk_rows && e->ready && !got) got = e;
  }
  if (!whole && !got && free_slot && qe_settings_u64("QE_RESIDENT_WHOLE", 1) && res_room_locked(rows * col->width + 64)) {
    const U64 nchunks = (rows + chunk_rows - 1) / chunk_rows;
    U8* state = (U8*)calloc(nchunks, 1);
    if (state && (whole = res_new_entry_locked(free_slot, col->uid, QE_WHOLE_COLUMN, chunk_rows, rows * col->width + 64))) {
      whole->nchunks = nchunks;
      whole->state = state;
      whole->uses = 0;
      free_slot = 0;
    } else {
      free(state);
    }
  }
  /* a filled column serves any chunk size; a partly filled one only the
   * chunk size its states were kept in */
  if (whole && whole->ready) {
    GPU_Buffer* v = gpu_hip_buffer_view(whole->buffer, c0 * col->width, bytes);
    if (v) {
      whole->pins++;
      whole->last_use = ++g_res.tick;
      if (chunk == 0) whole->uses++;
      g_res.hit_bytes += bytes;
      *view = v;
      res_unlock();
      return whole;
    }
  } else if (whole && whole->chunk_rows == chunk_rows && chunk < whole->nchunks && whole->state[chunk] != 1) {
    GPU_Buffer* v = gpu_hip_buffer_view(whole->buffer, c0 * col->width, bytes);
    if (v) {
      whole->pins++;
      whole->last_use = ++g_res.tick;
      if (chunk == 0) whole->uses++;
      if (whole->state[chunk] == 2) {
        g_res.hit_bytes += bytes;
      } else {
        whole->state[chunk] = 1;
        g_res.miss_bytes += bytes;
        *needs_copy = 1;
      }
      *view = v;
      res_unlock();
      return whole;
    }
  }
  if (got) {
    got->last_use = ++g_res.tick;
    got->uses++;
    got->pins++;
    g_res.hit_bytes += bytes;
    res_unlock();
    return got;
  }
  if (!whole && free_slot && res_room_locked(bytes)) {
    got = res_new_entry_locked(free_slot, col->uid, chunk, chunk_rows, bytes);
    if (got) {
      got->pins = 1;
      g_res.miss_bytes += bytes;
      *needs_copy = 1;
    }
  }
  res_unlock();
  return got;
}

/* `filled`: the chunk's rows are in the buffer; a chunk entry never filled
 * is freed, a whole-column entry's chunk goes back to empty */
static void res_unpin(QE_ResidentEntry* e, U64 chunk, B32 filled)
{
  if (!e) return;
  res_lock();
  if (e->used) {
    if (e->pins) e->pins--;
    if (e->chunk == QE_WHOLE_COLUMN) {
      if (!e->ready && chunk < e->nchunks && e->state[chunk] == 1) {
        e->state[chunk] = filled ? 2 : 0;
        if (filled) e->nfilled++;
        e->ready = e->nfilled == e->nchunks;
      }
    } else if (filled) {
      e->ready = 1;
    } else if (!e->ready && e->pins == 0) {
      res_free_entry(e);
    }
  }
  res_unlock();
}

/* every column of `cols` whole and filled in the cache: pin them and put
 * their buffers in bufs[column] */
static B32 res_acquire_whole(const QE_Table* t, const U32* cols, U32 n, QE_ResidentEntry** held, GPU_Buffer** bufs)
{
  if (qe_settings_u64("QE_RESIDENT", 1) == 0 || n == 0) return 0;
  res_lock();
  for (U32 i = 0; i < n; i++) {
    held[i] = 0;
    for (U32 k = 0; k < QE_RESIDENT_MAX_ENTRIES && !held[i]; k++) {
      QE_ResidentEntry* e = &g_res.entries[k];
      if (e->used && e->uid == t->columns[cols[i]].uid && e->chunk == QE_WHOLE_COLUMN && e->ready) held[i] = e;
    }
    if (!held[i]) {
      res_unlock();
      return 0;
    }
  }
  for (U32 i = 0; i < n; i++) {
    QE_ResidentEntry* e = held[i];
    e->pins++;
    e->uses++;
    e->last_use = ++g_res.tick;
    g_res.hit_bytes += e->bytes;
    bufs[cols[i]] = e->buffer;
  }
  res_unlock();
  return 1;
}

static void res_release_whole(QE_ResidentEntry** held, U32 n)
{
  res_lock();
  for (U32 i = 0; i < n; i++)
    if (held[i] && held[i]->pins) held[i]->pins--;
  res_unlock();
}

void qe_resident_drop_column(U64 uid)
{
  if (!uid) return;
  res_lock();
  for (U32 i = 0; i < QE_RESIDENT_MAX_ENTRIES; i++) {
    QE_ResidentEntry* e = &g_res.entries[i];
    if (e->used && e->uid == uid && e->pins == 0) res_free_entry(e);
  }
  res_unlock();
}

void qe_resident_clear(void)
{
  res_lock();
  for (U32 i = 0; i < QE_RESIDENT_MAX_ENTRIES; i++) {
    QE_ResidentEntry* e = &g_res.entries[i];
    if (e->used && e->pins == 0) res_free_entry(e);
  }
  res_unlock();
}

void qe_resident_stats(QE_ResidentStats* out)
{
  if (!out) return;
  res_lock();
  out->hit_bytes = g_res.hit_bytes;
  out->miss_bytes = g_res.miss_bytes;
  out->evicted_bytes = g_res.evicted_bytes;
  out->bytes_in_use = g_res.bytes_in_use;
  out->entries = g_res.entry_count;
  out->budget_bytes = res_budget_locked();
  res_unlock();
}

/* ------------------------------------------------------------------ */
/* host-backed columns                                                  */
/* ------------------------------------------------------------------ */
B32 qe_require_device(const QE_Table* t, const char* op)
{
  if (t && t->host_backed) {
    for (U32 c = 0; c < t->column_count; c++) {
      if (!t->columns[c].buffer) {
        qe_set_error("%s: the host-backed table's rows cannot be mapped for the device", op);
        return 0;
      }
    }
  }
  return 1;
}

B32 qe_refuse_host(const QE_Table* t, const char* op)
{
  if (t && t->host_backed) {
    qe_set_error("%s: the table is host-backed (its rows are read-only host memory)", op);
    return 0;
  }
  return 1;
}

/* zone maps of a host column, built on the device chunk by chunk with the
 * same kernel as a resident column's (the rows pass through one staging
 * buffer; nothing is cached) */
static B32 qe_host_zonemap(QE_Table* t, QE_Column* col)
{
  U64 rows = t->row_count;
  U64 zones = (rows + (1ull << QE_ZONE_SHIFT) - 1) >> QE_ZONE_SHIFT;
  col->zone_count = zones;
  col->zonemap_valid = 1;
  if (!zones) return 1;
  col->zone_min = (F64*)malloc(sizeof(F64) * zones);
  col->zone_max = (F64*)malloc(sizeof(F64) * zones);
  U64 chunk = qe_stream_chunk_rows(col->width);
  if (chunk > rows) chunk = (rows + (1ull << QE_ZONE_SHIFT) - 1) & ~((1ull << QE_ZONE_SHIFT) - 1);
  U64 czones = chunk >> QE_ZONE_SHIFT;
  GPU_Buffer* stage = gpu_buffer_alloc(chunk * col->width + 64, GPU_BufferFlag_ReadWrite | GPU_BufferFlag_DeviceLocal, 0);
  GPU_Buffer* bmin = qe_scratch("zonemap_min", czones * 8);
  GPU_Buffer* bmax = qe_scratch("zonemap_max", czones * 8);
  GPU_Kernel* k = qe_kernel("zonemap_build");
  B32 ok = stage && bmin && bmax && k && col->zone_min && col->zone_max;
  for (U64 c0 = 0; ok && c0 < rows; c0 += chunk) {
    U64 n = rows - c0 < chunk ? rows - c0 : chunk;
    U64 nz = (n + (1ull << QE_ZONE_SHIFT) - 1) >> QE_ZONE_SHIFT;
    U64 z0 = c0 >> QE_ZONE_SHIFT;
    gpu_kernel_set_arg_buffer(k, 6, stage);
    gpu_kernel_set_arg_buffer(k, 1, bmin);
    gpu_kernel_set_arg_buffer(k, 2, bmax);
    gpu_kernel_set_arg_u64(k, 0, n);
    qe_kernel_set_word(k, QK_W_ZTYPE, col->type);
    qe_kernel_set_word(k, QK_W_ZSHIFT, QE_ZONE_SHIFT);
    U64 blocks = nz < 65536 ? nz : 65536;
    GPU_Batch* b = gpu_batch_begin(n * col->width, nz * 16);
    gpu_batch_buffer_write(b, stage, (U8*)col->host_data + c0 * col->width, n * col->width);
    gpu_batch_kernel_execute(b, k, (U32)(blocks * QE_WG), QE_WG);
    gpu_batch_buffer_read(b, bmin, col->zone_min + z0, nz * 8);
    gpu_batch_buffer_read(b, bmax, col->zone_max + z0, nz * 8);
    ok = gpu_batch_end(b);
  }
  if (stage) gpu_buffer_release(stage);
  qe_scratch_release(bmin);
  qe_scratch_release(bmax);
  if (!ok) {
    qe_set_error("qe_table_add_host_column: zone maps: %s", gpu_hip_last_error());
    return 0;
  }
  F64 mn = col->zone_min[0], mx = col->zone_max[0];
  for (U64 z = 1; z < zones; z++) {
    if (col->zone_min[z] < mn) mn = col->zone_min[z];
    if (col->zone_max[z] > mx) mx = col->zone_max[z];
  }
  col->min_value = mn;
  col->max_value = mx;
  qe_zonemap_build_super(col);
  return 1;
}

S32 qe_table_add_host_column(QE_Table* table, const char* name, U32 type, const void* data)
{
  if (!table || table->column_count >= QE_MAX_COLUMNS) {
    qe_set_error("qe_table_add_host_column: table full");
    return -1;
  }
  if (table->column_count && !table->host_backed) {
    qe_set_error("qe_table_add_host_column: the table already holds device columns");
    return -1;
  }
  U32 width = qe_type_width(type);
  if (!width) {
    qe_set_error("qe_table_add_host_column: unsupported type %u", type);
    return -1;
  }
  if (!data && table->row_count) {
    qe_set_error("qe_table_add_host_column: no rows given");
    return -1;
  }
  U64 bytes = table->row_count * width;
  void* host = gpu_hip_host_alloc(bytes + 64);
  if (!host) {
    qe_set_error("qe_table_add_host_column: pinned allocation of %llu bytes failed: %s", (unsigned long long)bytes,
                 gpu_hip_last_error());
    return -1;
  }
  if (bytes) memcpy(host, data, bytes);
  memset((U8*)host + bytes, 0, 64);
  QE_Column* col = &table->columns[table->column_count];
  memset(col, 0, sizeof(*col));
  snprintf(col->name, sizeof(col->name), "%s", name ? name : "");
  col->type = type;
  col->width = width;
  col->host_data = host;
  /* the rows in place for the operators that do not stream (0 if the
   * device cannot map them: those operators then refuse the table) */
  col->buffer = gpu_hip_buffer_wrap_host(host, bytes + 64);
  res_lock();
  col->uid = g_next_uid++;
  res_unlock();
  table->host_backed = 1;
  S32 index = (S32)table->column_count++;
  if (!qe_host_zonemap(table, col)) {
    table->column_count--;
    free(col->zone_min);
    free(col->zone_max);
    free(col->super_zone);
    if (col->buffer) gpu_buffer_release(col->buffer);
    gpu_hip_host_free(host);
    memset(col, 0, sizeof(*col));
    if (table->column_count == 0) table->host_backed = 0;
    return -1;
  }
  return index;
}

/* ------------------------------------------------------------------ */
/* the chunk pipeline                                                   */
/* ------------------------------------------------------------------ */
static __thread QE_StreamStats qe_t_stream_stats;

void qe_stream_last_stats(QE_StreamStats* out)
{
  if (out) *out = qe_t_stream_stats;
}

typedef struct QE_ChunkSlot {
  U64 chunk, c0, n;
  B32 copying; /* copies were issued on this slot */
  QE_ResidentEntry* held[QE_MAX_COLUMNS];
  GPU_Buffer* own[QE_MAX_COLUMNS]; /* buffers outside the cache, or views of a whole-column entry */
  GPU_Buffer* buf[QE_MAX_COLUMNS];
} QE_ChunkSlot;

typedef B32 (*QE_ChunkFn)(void* ctx, QE_Table* view, U64 c0);

typedef struct QE_Pipeline {
  QE_Table* t;
  U32 cols[QE_MAX_COLUMNS];
  U32 col_count;
  U64 chunk_rows;
  GPU_Stream* copy;
  U32 nslots;
  QE_ChunkSlot slot[QE_STREAM_MAX_SLOTS];
  QE_Table* view;
} QE_Pipeline;

static void qe_slot_release(QE_Pipeline* p, QE_ChunkSlot* s, B32 filled)
{
  for (U32 i = 0; i < p->col_count; i++) {
    U32 c = p->cols[i];
    if (s->held[c]) res_unpin(s->held[c], s->chunk, filled);
    if (s->own[c]) gpu_buffer_release(s->own[c]);
    s->held[c] = 0;
    s->own[c] = 0;
    s->buf[c] = 0;
  }
  s->copying = 0;
}

/* acquire chunk buffers for every column read and start the copies of
 * those not resident yet */
static B32 qe_slot_prepare(QE_Pipeline* p, U32 si, U64 chunk)
{
  QE_ChunkSlot* s = &p->slot[si];
  s->chunk = chunk;
  s->c0 = chunk * p->chunk_rows;
  s->n = p->t->row_count - s->c0 < p->chunk_rows ? p->t->row_count - s->c0 : p->chunk_rows;
  s->copying = 0;
  B32 all_hit = 1;
  for (U32 i = 0; i < p->col_count; i++) {
    QE_Column* col = &p->t->columns[p->cols[i]];
    U64 rows_bytes = s->n * col->width;
    /* 64 bytes of slack keep the kernels' 16-byte loads inside the buffer */
    U64 bytes = rows_bytes + 64;
    B32 needs_copy = 0;
    GPU_Buffer* v = 0;
    QE_ResidentEntry* e = res_acquire(col, p->t->row_count, chunk, p->chunk_rows, &needs_copy, &v);
    GPU_Buffer* b = v ? v : (e ? e->buffer : 0);
    if (v) s->own[p->cols[i]] = v;
    if (!b) {
      needs_copy = 1;
      b = gpu_buffer_alloc(bytes, GPU_BufferFlag_ReadWrite | GPU_BufferFlag_DeviceLocal, 0);
      if (!b) {
        qe_set_error("streamed scan: chunk buffer of %llu bytes: %s", (unsigned long long)bytes, gpu_hip_last_error());
        return 0;
      }
      s->own[p->cols[i]] = b;
      qe_t_stream_stats.transient_chunks++;
    }
    s->held[p->cols[i]] = e;
    s->buf[p->cols[i]] = b;
    if (needs_copy) {
      all_hit = 0;
      if (!s->copying) {
        if (!gpu_hip_stream_slot_begin(p->copy, si, 0, 0)) return 0;
        s->copying = 1;
      }
      if (!gpu_hip_stream_copy(p->copy, si, b, (U8*)col->host_data + s->c0 * col->width, rows_bytes)) {
        qe_set_error("streamed scan: copy: %s", gpu_hip_last_error());
        return 0;
      }
      qe_t_stream_stats.bytes_copied += rows_bytes;
    }
  }
  if (all_hit) qe_t_stream_stats.chunks_resident++;
  if (s->copying && !gpu_hip_stream_submit(p->copy, si)) return 0;
  return 1;
}

/* the chunk as a table: its rows, zone maps and buffers */
static void qe_view_of(QE_Pipeline* p, QE_ChunkSlot* s)
{
  QE_Table* v = p->view;
  const QE_Table* t = p->t;
  v->row_count = s->n;
  v->column_count = t->column_count;
  v->host_backed = 0;
  U64 z0 = s->c0 >> QE_ZONE_SHIFT;
  U64 nz = (s->n + (1ull << QE_ZONE_SHIFT) - 1) >> QE_ZONE_SHIFT;
  for (U32 c = 0; c < t->column_count; c++) {
    QE_Column* d = &v->columns[c];
    *d = t->columns[c];
    d->buffer = s->buf[c];
    d->host_data = 0;
    d->uid = 0;
    d->enc = 0;
    d->enc_base = 0;
    d->enc_buffer = 0;
    d->zone_count = nz;
    d->zone_min = t->columns[c].zone_min ? t->columns[c].zone_min + z0 : 0;
    d->zone_max = t->columns[c].zone_max ? t->columns[c].zone_max + z0 : 0;
    /* the column's zone groups serve the chunk when it starts on a group
     * and ends on one or at the column's end */
    const U64 gz = 1ull << QE_SUPER_SHIFT;
    const B32 aligned = (z0 % gz) == 0 && ((nz % gz) == 0 || s->c0 + s->n == t->row_count);
    d->super_zone = aligned && t->columns[c].super_zone ? t->columns[c].super_zone + 4 * (z0 / gz) : 0;
    d->super_count = d->super_zone ? (nz + gz - 1) / gz : 0;
  }
}

static void qe_mark_column(QE_Pipeline* p, U32 c)
{
  for (U32 i = 0; i < p->col_count; i++)
    if (p->cols[i] == c) return;
  p->cols[p->col_count++] = c;
}

static B32 qe_mark_where(QE_Pipeline* p, const QE_IRNode* where, U32 count)
{
  for (U32 i = 0; i < count; i++) {
    if (where[i].type != QE_IR_COLUMN) continue;
    if (where[i].column >= p->t->column_count) {
      qe_set_error("streamed scan: unknown column %u", where[i].column);
      return 0;
    }
    qe_mark_column(p, where[i].column);
  }
  return 1;
}

/* run fn over every chunk a zone of which survives `leaves`, copies of
 * chunk i+1 overlapping the kernels of chunk i */
static B32 qe_pipeline_run(QE_Pipeline* p, const QE_Leaf* leaves, U32 leaf_count, QE_ChunkFn fn, void* ctx)
{
  QE_Table* t = p->t;
  U32 widest = 0;
  for (U32 i = 0; i < p->col_count; i++)
    if (t->columns[p->cols[i]].width > widest) widest = t->columns[p->cols[i]].width;
  p->chunk_rows = qe_stream_chunk_rows(widest);
  p->nslots = qe_stream_slots();
  U64 chunks = (t->row_count + p->chunk_rows - 1) / p->chunk_rows;
  /* warm: every column read is whole in the cache; one pass over them as
   * over a device-resident table */
  {
    QE_ResidentEntry* held[QE_MAX_COLUMNS];
    GPU_Buffer* bufs[QE_MAX_COLUMNS];
    memset(bufs, 0, sizeof(bufs));
    if (res_acquire_whole(t, p->cols, p->col_count, held, bufs)) {
      QE_Table* v = (QE_Table*)malloc(sizeof(QE_Table));
      B32 ok = v != 0;
      if (ok) {
        *v = *t;
        v->host_backed = 0;
        for (U32 c = 0; c < t->column_count; c++) {
          QE_Column* d = &v->columns[c];
          d->buffer = bufs[c];
          d->host_data = 0;
          d->uid = 0;
        }
        ok = fn(ctx, v, 0);
        free(v);
      } else {
        qe_set_error("streamed scan: out of host memory");
      }
      res_release_whole(held, p->col_count);
      qe_t_stream_stats.chunks = chunks;
      qe_t_stream_stats.chunks_resident = chunks;
      qe_t_stream_stats.resident_whole = 1;
      return ok;
    }
  }
  QE_Prune prune;
  qe_zonemap_prune(t, leaves, leaf_count, &prune);
  U64* run = (U64*)malloc(sizeof(U64) * (chunks ? chunks : 1));
  p->view = (QE_Table*)calloc(1, sizeof(QE_Table));
  if (!run || !p->view) {
    free(run);
    free(p->view);
    qe_prune_free(&prune);
    qe_set_error("streamed scan: out of host memory");
    return 0;
  }
  U64 nrun = 0;
  U64 zpc = p->chunk_rows >> QE_ZONE_SHIFT;
  for (U64 c = 0; c < chunks; c++) {
    B32 live = prune.zones_left == prune.zones_total;
    if (!live && prune.zones_left) {
      U64 z1 = (c + 1) * zpc < prune.zones_total ? (c + 1) * zpc : prune.zones_total;
      for (U64 z = c * zpc; z < z1 && !live; z++) live = !prune.skip_bits || !((prune.skip_bits[z >> 5] >> (z & 31)) & 1);
    }
    if (live) run[nrun++] = c;
  }
  qe_prune_free(&prune);
  qe_t_stream_stats.chunks = chunks;
  qe_t_stream_stats.chunks_skipped = chunks - nrun;

  B32 ok = 1;
  const U32 S = p->nslots;
  p->copy = nrun ? gpu_hip_stream_alloc(S) : 0;
  if (nrun && !p->copy) {
    qe_set_error("streamed scan: no copy stream: %s", gpu_hip_last_error());
    ok = 0;
  }
  /* chunk i runs in slot i % S; the copies of chunks i + 1 .. i + S - 1
   * are in flight meanwhile */
  for (U64 i = 0; ok && i + 1 < S && i < nrun; i++) ok = qe_slot_prepare(p, (U32)i, run[i]);
  for (U64 i = 0; ok && i < nrun; i++) {
    U32 si = (U32)(i % S);
    if (i + S - 1 < nrun) ok = qe_slot_prepare(p, (U32)((i + S - 1) % S), run[i + S - 1]);
    QE_ChunkSlot* s = &p->slot[si];
    if (ok && s->copying) {
      U64 w0 = qe_now_ns();
      ok = gpu_hip_stream_wait(p->copy, si);
      qe_t_stream_stats.copy_wait_ns += qe_now_ns() - w0;
      s->copying = 0;
      if (!ok) qe_set_error("streamed scan: copy: %s", gpu_hip_last_error());
    }
    if (ok) {
      qe_view_of(p, s);
      ok = fn(ctx, p->view, s->c0);
    }
    qe_slot_release(p, s, ok);
    qe_t_stream_stats.chunks_run += ok;
  }
  /* after a failure, copies still in flight finish before their buffers go */
  if (p->copy) {
    for (U32 si = 0; si < S; si++) {
      if (p->slot[si].copying) gpu_hip_stream_wait(p->copy, si);
      qe_slot_release(p, &p->slot[si], 0);
    }
    gpu_hip_stream_release(p->copy);
  }
  free(run);
  free(p->view);
  p->view = 0;
  return ok;
}

static void qe_trace_add(QE_Trace* sum, const QE_Trace* part)
{
  if (!sum) return;
  if (!sum->path) {
    sum->path = part->path;
    memcpy(sum->kernel, part->kernel, sizeof(sum->kernel));
  }
  sum->kernel_count += part->kernel_count;
  sum->rows_scanned += part->rows_scanned;
  sum->bytes_read += part->bytes_read;
  sum->gpu_kernel_ns += part->gpu_kernel_ns;
  sum->compile_ns += part->compile_ns;
}

/* ------------------------------------------------------------------ */
/* streamed scan (qe_scan_filter_streamed, query_exec.c:2023-2491)      */
/* ------------------------------------------------------------------ */
typedef struct QE_ScanCtx {
  const QE_IRNode* where;
  U32 node_count;
  B32 count_only;
  U64* ids;
  U64 cap;
  U64 count;
  QE_Trace* trace;
} QE_ScanCtx;

static B32 qe_scan_chunk(void* vctx, QE_Table* view, U64 c0)
{
  QE_ScanCtx* x = (QE_ScanCtx*)vctx;
  U64 room = x->count < x->cap ? x->cap - x->count : 0;
  B32 only = x->count_only || room == 0 || !x->ids;
  U64 n = 0;
  QE_Trace part;
  memset(&part, 0, sizeof(part));
  if (!qe_scan_filter(view, x->where, x->node_count, only, only ? 0 : x->ids + x->count, room, &n, &part)) return 0;
  if (!only) {
    U64 wrote = n < room ? n : room;
    for (U64 i = 0; i < wrote; i++) x->ids[x->count + i] += c0;
  }
  x->count += n;
  qe_trace_add(x->trace, &part);
  return 1;
}

B32 qe_stream_scan(QE_Table* table, const QE_IRNode* where, U32 node_count, B32 count_only, U64* out_ids, U64 cap,
                   U64* out_count, QE_Trace* trace)
{
  U64 t0 = qe_now_ns();
  memset(&qe_t_stream_stats, 0, sizeof(qe_t_stream_stats));
  if (trace) {
    memset(trace, 0, sizeof(*trace));
    trace->rows_before = table->row_count;
  }
  QE_Pipeline p;
  memset(&p, 0, sizeof(p));
  p.t = table;
  if (!qe_mark_where(&p, where, where ? node_count : 0)) return 0;
  QE_Leaf leaves[32];
  U32 leaf_count = where && node_count ? qe_collect_prune_leaves(where, node_count, leaves, 32) : 0;
  QE_ScanCtx x;
  memset(&x, 0, sizeof(x));
  x.where = where;
  x.node_count = node_count;
  x.count_only = count_only;
  x.ids = out_ids;
  x.cap = out_ids ? cap : 0;
  x.trace = trace;
  if (!qe_pipeline_run(&p, leaves, leaf_count, qe_scan_chunk, &x)) return 0;
  *out_count = x.count;
  qe_t_stream_stats.wall_ns = qe_now_ns() - t0;
  if (trace) {
    if (!trace->path) trace->path = QE_ScanPath_ZoneEmpty;
    trace->rows_after = x.count;
    trace->wall_ns = qe_t_stream_stats.wall_ns;
  }
  return 1;
}

/* ------------------------------------------------------------------ */
/* streamed aggregate: per-chunk partial groups merged per key          */
/* (qe_aggregate_global_streamed :8842-9110 and the dense streamed      */
/* path :7984-8049 fold per-chunk partials on the host the same way)    */
/* ------------------------------------------------------------------ */
#define QE_STREAM_MAX_PARTS (1 + 3 * QE_AGG_MAX_ARGS)

typedef struct QE_Group {
  U64 key_bits[QE_AGG_MAX_GROUP_COLS]; /* F64 bits of the key values, 0 past K */
  U64 count;
  U64 part[QE_STREAM_MAX_PARTS]; /* S64 sum, or F64 bits of sum/min/max */
} QE_Group;

typedef struct QE_AggCtx {
  const QE_IRNode* where;
  U32 where_count;
  const S32* group_cols;
  U32 key_count;
  /* the partial expressions: 0 is COUNT(*), then per argument column its
   * SUM, MIN and MAX as the query needs them */
  U32 part_count;
  U32 part_func[QE_STREAM_MAX_PARTS];
  S32 part_col[QE_STREAM_MAX_PARTS];
  B32 part_int[QE_STREAM_MAX_PARTS]; /* SUM over a 32-bit integer column */
  /* groups: an open-addressed index over key bits */
  QE_Group* groups;
  U64 group_count, group_cap;
  U64* slots; /* index + 1, 0 empty */
  U64 slot_mask;
  /* chunk results */
  U64 out_cap;
  F64* keys;
  U64* cells;
  QE_Trace* trace;
} QE_AggCtx;

static U64 qe_mix64(U64 x)
{
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  return x ^ (x >> 33);
}

static U64 qe_key_hash(const U64* key_bits)
{
  U64 h = 0;
  for (U32 k = 0; k < QE_AGG_MAX_GROUP_COLS; k++) h = qe_mix64(h ^ (key_bits[k] + 0x9E3779B97F4A7C15ull * k));
  return h;
}

static QE_Group* qe_group_find(QE_AggCtx* x, const U64* key_bits)
{
  if ((x->group_count + 1) * 2 > x->slot_mask + 1) {
    U64 nslots = (x->slot_mask + 1) * 2;
    U64* ns = (U64*)calloc(nslots, sizeof(U64));
    if (!ns) return 0;
    for (U64 i = 0; i < x->group_count; i++) {
      U64 h = qe_key_hash(x->groups[i].key_bits) & (nslots - 1);
      while (ns[h]) h = (h + 1) & (nslots - 1);
      ns[h] = i + 1;
    }
    free(x->slots);
    x->slots = ns;
    x->slot_mask = nslots - 1;
  }
  U64 h = qe_key_hash(key_bits) & x->slot_mask;
  while (x->slots[h]) {
    QE_Group* g = &x->groups[x->slots[h] - 1];
    if (!memcmp(g->key_bits, key_bits, sizeof(g->key_bits))) return g;
    h = (h + 1) & x->slot_mask;
  }
  if (x->group_count == x->group_cap) {
    U64 nc = x->group_cap ? x->group_cap * 2 : 1024;
    QE_Group* ng = (QE_Group*)realloc(x->groups, sizeof(QE_Group) * nc);
    if (!ng) return 0;
    x->groups = ng;
    x->group_cap = nc;
  }
  QE_Group* g = &x->groups[x->group_count];
  memset(g, 0, sizeof(*g));
  memcpy(g->key_bits, key_bits, sizeof(g->key_bits));
  x->slots[h] = ++x->group_count;
  return g;
}

static void qe_part_merge(QE_AggCtx* x, QE_Group* g, U32 i, U64 cell, B32 first)
{
  F64 a, b;
  switch (x->part_func[i]) {
    case QE_AGG_FUNC_SUM:
      if (x->part_int[i]) {
        g->part[i] = (U64)((S64)g->part[i] + (S64)cell);
      } else {
        memcpy(&a, &g->part[i], 8);
        memcpy(&b, &cell, 8);
        a = first ? b : a + b;
        memcpy(&g->part[i], &a, 8);
      }
      break;
    case QE_AGG_FUNC_MIN:
    case QE_AGG_FUNC_MAX:
      if (first) {
        g->part[i] = cell;
      } else {
        memcpy(&a, &g->part[i], 8);
        memcpy(&b, &cell, 8);
        if (x->part_func[i] == QE_AGG_FUNC_MIN ? b < a : b > a) g->part[i] = cell;
      }
      break;
    default: break;
  }
}

static B32 qe_agg_chunk(void* vctx, QE_Table* view, U64 c0)
{
  (void)c0;
  QE_AggCtx* x = (QE_AggCtx*)vctx;
  /* the partials in packs of at most QE_AGG_MAX_EXPRS, each led by COUNT(*) */
  for (U32 p0 = 1; p0 < x->part_count || p0 == 1; p0 += QE_AGG_MAX_EXPRS - 1) {
    U32 funcs[QE_AGG_MAX_EXPRS];
    S32 cols[QE_AGG_MAX_EXPRS];
    U32 map[QE_AGG_MAX_EXPRS];
    U32 n = 1;
    funcs[0] = QE_AGG_FUNC_COUNT;
    cols[0] = -1;
    map[0] = 0;
    for (U32 i = p0; i < x->part_count && n < QE_AGG_MAX_EXPRS; i++, n++) {
      funcs[n] = x->part_func[i];
      cols[n] = x->part_col[i];
      map[n] = i;
    }
    U64 groups = 0;
    for (;;) {
      QE_Trace part;
      memset(&part, 0, sizeof(part));
      qe_t_raw_int_sums = 1;
      B32 ok = qe_aggregate(view, x->where, x->where_count, x->group_cols, x->key_count, funcs, cols, n, x->out_cap,
                            &groups, x->keys, x->cells, &part);
      qe_t_raw_int_sums = 0;
      if (!ok) return 0;
      qe_trace_add(x->trace, &part);
      if (groups <= x->out_cap) break;
      U64 nc = groups + groups / 4;
      F64* nk = (F64*)realloc(x->keys, sizeof(F64) * nc * (x->key_count ? x->key_count : 1));
      if (nk) x->keys = nk;
      U64* ncl = (U64*)realloc(x->cells, sizeof(U64) * nc * QE_AGG_MAX_EXPRS);
      if (ncl) x->cells = ncl;
      if (!nk || !ncl) {
        qe_set_error("streamed aggregate: out of host memory");
        return 0;
      }
      x->out_cap = nc;
    }
    for (U64 g = 0; g < groups; g++) {
      U64 cnt = x->cells[g * n];
      if (cnt == 0) continue; /* an ungrouped chunk with no row matching */
      U64 key_bits[QE_AGG_MAX_GROUP_COLS] = {0, 0, 0, 0};
      for (U32 k = 0; k < x->key_count; k++) {
        F64 v = x->keys[g * x->key_count + k];
        if (v == 0.0) v = 0.0; /* -0.0 and 0.0 are one key */
        memcpy(&key_bits[k], &v, 8);
      }
      QE_Group* grp = qe_group_find(x, key_bits);
      if (!grp) {
        qe_set_error("streamed aggregate: out of host memory");
        return 0;
      }
      /* the COUNT of the first pack is the group's (later packs see the
       * same rows); a group whose count is this chunk's is new */
      if (p0 == 1) grp->count += cnt;
      B32 first = grp->count == cnt;
      for (U32 e = 1; e < n; e++) qe_part_merge(x, grp, map[e], x->cells[g * n + e], first);
    }
    if (x->part_count <= 1) break;
  }
  return 1;
}

/* groups in ascending key order, key 0 first (unused key words are 0) */
static int qe_key_order(const void* a, const void* b)
{
  for (U32 k = 0; k < QE_AGG_MAX_GROUP_COLS; k++) {
    F64 ka, kb;
    const U64 ba = ((const QE_Group*)a)->key_bits[k], bb = ((const QE_Group*)b)->key_bits[k];
    memcpy(&ka, &ba, 8);
    memcpy(&kb, &bb, 8);
    if (ka < kb) return -1;
    if (ka > kb) return 1;
    if (ba != bb) return ba < bb ? -1 : 1;
  }
  return 0;
}

static B32 qe_type_int32_col(U32 type) { return type == CQL_ColumnType_U32 || type == CQL_ColumnType_I32; }

B32 qe_stream_aggregate(QE_Table* table, const QE_IRNode* where, U32 where_count, const S32* group_cols,
                        U32 group_count, const U32* funcs, const S32* arg_cols, U32 expr_count, U64 cap,
                        U64* out_group_count, F64* out_keys, U64* out_cells, QE_Trace* trace)
{
  U64 t0 = qe_now_ns();
  memset(&qe_t_stream_stats, 0, sizeof(qe_t_stream_stats));
  if (trace) {
    memset(trace, 0, sizeof(*trace));
    trace->rows_before = table->row_count;
  }
  QE_AggPlan plan;
  if (!qe_agg_plan(table, funcs, arg_cols, expr_count, &plan)) return 0;
  QE_GroupKeys gk;
  if (!qe_group_keys(table, group_cols, group_count, &gk)) return 0;
  const B32 grouped = group_count > 0;
  QE_AggCtx x;
  memset(&x, 0, sizeof(x));
  x.where = where;
  x.where_count = where_count;
  x.group_cols = group_cols;
  x.key_count = group_count;
  x.trace = trace;
  x.part_count = 1;
  x.part_func[0] = QE_AGG_FUNC_COUNT;
  x.part_col[0] = -1;
  /* per expression: where its value comes from among the partials */
  U32 src[QE_AGG_MAX_EXPRS];
  for (U32 e = 0; e < expr_count; e++) {
    src[e] = 0;
    S32 a = plan.expr_arg[e];
    if (a < 0) continue;
    U32 want = plan.funcs[e] == QE_AGG_FUNC_AVG ? QE_AGG_FUNC_SUM : plan.funcs[e];
    S32 col = (S32)plan.arg_col[a];
    U32 i = 1;
    while (i < x.part_count && !(x.part_func[i] == want && x.part_col[i] == col)) i++;
    if (i == x.part_count) {
      x.part_func[i] = want;
      x.part_col[i] = col;
      x.part_int[i] = want == QE_AGG_FUNC_SUM && qe_type_int32_col(table->columns[col].type);
      x.part_count++;
    }
    src[e] = i;
  }
  QE_Pipeline p;
  memset(&p, 0, sizeof(p));
  p.t = table;
  B32 ok = qe_mark_where(&p, where, where ? where_count : 0);
  for (U32 k = 0; ok && k < group_count; k++) qe_mark_column(&p, (U32)group_cols[k]);
  for (U32 a = 0; ok && a < plan.arg_count; a++) qe_mark_column(&p, plan.arg_col[a]);
  x.slots = (U64*)calloc(1024, sizeof(U64));
  x.slot_mask = 1023;
  x.out_cap = grouped ? 65536 : 1;
  x.keys = (F64*)malloc(sizeof(F64) * x.out_cap * (group_count ? group_count : 1));
  x.cells = (U64*)malloc(sizeof(U64) * x.out_cap * QE_AGG_MAX_EXPRS);
  if (ok && (!x.slots || !x.keys || !x.cells)) {
    qe_set_error("streamed aggregate: out of host memory");
    ok = 0;
  }
  if (ok) {
    QE_Leaf leaves[32];
    U32 leaf_count = where && where_count ? qe_collect_prune_leaves(where, where_count, leaves, 32) : 0;
    ok = qe_pipeline_run(&p, leaves, leaf_count, qe_agg_chunk, &x);
  }
  if (ok) {
    if (grouped) {
      qsort(x.groups, x.group_count, sizeof(QE_Group), qe_key_order);
      *out_group_count = x.group_count;
    } else {
      /* one row, zeros over no rows (query_exec.c:10335-10346) */
      if (x.group_count == 0) {
        const U64 none[QE_AGG_MAX_GROUP_COLS] = {0, 0, 0, 0};
        QE_Group* g = qe_group_find(&x, none);
        if (!g) ok = 0;
      }
      *out_group_count = 1;
    }
  }
  if (ok) {
    U64 rows = x.group_count < cap ? x.group_count : cap;
    for (U64 g = 0; g < rows; g++) {
      const QE_Group* grp = &x.groups[g];
      if (out_keys)
        for (U32 k = 0; k < group_count; k++) memcpy(&out_keys[g * group_count + k], &grp->key_bits[k], 8);
      for (U32 e = 0; e < expr_count; e++) {
        U32 f = plan.funcs[e], i = src[e];
        U64* cell = &out_cells[g * expr_count + e];
        if (f == QE_AGG_FUNC_COUNT || grp->count == 0) {
          *cell = qe_agg_cell(f, grp->count, 0.0, 0.0, 0.0);
        } else if (x.part_int[i]) {
          *cell = qe_agg_cell_int(f, grp->count, (S64)grp->part[i], 0.0, 0.0, 0);
        } else {
          F64 v;
          memcpy(&v, &grp->part[i], 8);
          *cell = qe_agg_cell(f, grp->count, v, v, v);
        }
      }
    }
  }
  free(x.groups);
  free(x.slots);
  free(x.keys);
  free(x.cells);
  qe_t_stream_stats.wall_ns = qe_now_ns() - t0;
  if (ok && trace) {
    if (!trace->path) trace->path = grouped ? QE_AggPath_Dense : QE_AggPath_FusedGlobal;
    trace->rows_after = *out_group_count;
    trace->wall_ns = qe_t_stream_stats.wall_ns;
  }
  return ok;
}
