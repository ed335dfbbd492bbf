/*
 * qe_serve.c - served request loop with continuous batching.
 *
 * The reference serves its Postgres-wire clients one query at a time: each
 * connection's query runs under the global g_query_exec_mutex
 * (application.h:34, pg_server.c:200-250 and :643-760).  This loop sits
 * where that mutex is: requests are queued; worker threads (one per batch
 * in flight) each take the queue head together with every queued request
 * on the same table that still fits, up to max_batch queries and while the
 * device memory the admitted requests need stays within hbm_budget, and run
 * them as one qe_aggregate_batch on their own stream.  Requests that arrive
 * while batches are on the GPU wait in the queue and join the next batch a
 * worker forms, so a worker's host work and the other worker's kernels
 * overlap.  Tables stay resident; nothing is uploaded per request.
 * Latency is measured per request from submit to completion.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <time.h>

#include "qe_internal.h"

#define QE_SERVE_MAX_WORKERS 8

typedef struct QE_Request {
  QE_Table* table;
  U32 kind; /* QE_REQUEST_* */
  QE_BatchQuery* q;
  QE_JoinQuery* join;
  QE_ScanQuery* scan;
  U64 footprint;
  U64 t_submit, t_done;
  U32 state; /* 0 queued, 1 running, 2 done */
  struct QE_Request* next;
} QE_Request;

struct QE_Server {
  pthread_mutex_t lock;
  pthread_cond_t work_cv, done_cv;
  pthread_t threads[QE_SERVE_MAX_WORKERS];
  U32 workers, max_batch;
  U64 hbm_budget;
  QE_Request* head;
  QE_Request* tail;
  U64 queued;
  U64 inflight_bytes;
  B32 stop;
  /* stats since the last reset */
  U64 submitted, completed, failed, batches, batched_queries, joins, scans;
  U64 t_first_submit, t_last_done;
  U64 inflight_peak_bytes, queue_peak;
  U64* lat;
  U64 lat_count, lat_cap;
};

/* device memory a request needs while it runs: the accumulators of its
 * groups (dense range from the zone map, else up to one per row) in the
 * partial, result and output copies */
static U64 qe_serve_footprint(QE_Table* t, const QE_BatchQuery* q)
{
  U64 groups = 1;
  QE_GroupKeys keys;
  if (q->group_count && qe_group_keys(t, q->group_cols, q->group_count, &keys)) {
    groups = t->row_count ? t->row_count : 1;
    const U64 span = qe_group_keys_span(&keys);
    if (span && span < groups) groups = span;
  }
  return 65536 + groups * (1 + 3 * 4 + 2 * QE_AGG_MAX_GROUP_COLS) * 8 * 4;
}

/* a join's partition buffers (key and arguments of every fact row, in
 * partition order) and its dimension's table; a scan's id buffer and block
 * counts */
static U64 qe_serve_footprint_join(QE_Table* fact, const QE_JoinQuery* j)
{
  const U64 dim_rows = j->dim ? j->dim->row_count : 0;
  return 65536 + fact->row_count * 4 * (1 + (U64)j->expr_count) + dim_rows * 16 + j->cap * 8 * 8;
}

static U64 qe_serve_footprint_scan(QE_Table* t, const QE_ScanQuery* q)
{
  return 65536 + (q->count_only ? 0 : q->cap * 8) + t->row_count / 256;
}

static int qe_u64_cmp(const void* a, const void* b)
{
  U64 x = *(const U64*)a, y = *(const U64*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static void* qe_serve_worker(void* arg)
{
  QE_Server* s = (QE_Server*)arg;
  QE_Request** batch = (QE_Request**)calloc(s->max_batch, sizeof(QE_Request*));
  QE_BatchQuery* qs = (QE_BatchQuery*)calloc(s->max_batch, sizeof(QE_BatchQuery));
  QE_JoinQuery* jq = (QE_JoinQuery*)calloc(s->max_batch, sizeof(QE_JoinQuery));
  if (!batch || !qs || !jq) {
    free(batch);
    free(qs);
    free(jq);
    return 0;
  }
  pthread_mutex_lock(&s->lock);
  for (;;) {
    while (!s->stop && !s->head) pthread_cond_wait(&s->work_cv, &s->lock);
    if (!s->head && s->stop) break;
    /* admit: the head, then every queued request on its table that fits
     * the batch and the memory budget (the head always, when nothing else
     * is in flight, so an oversized request still runs) */
    QE_Table* table = s->head->table;
    const U32 kind = s->head->kind;
    U32 n = 0;
    QE_Request** link = &s->head;
    QE_Request* prev = 0;
    while (*link && n < (kind == QE_REQUEST_SCAN ? 1 : s->max_batch)) {
      QE_Request* r = *link;
      B32 fits = s->inflight_bytes + r->footprint <= s->hbm_budget || (n == 0 && s->inflight_bytes == 0);
      if (r->kind != kind) {
        /* scans run one per launch; aggregates batch with the aggregates
         * queued behind them, joins with the joins on the same fact table
         * (qe_aggregate_join_batch: identical ones once, a fact side shared) */
        if (n == 0) break;
        prev = r;
        link = &r->next;
        continue;
      }
      if (r->table == table && fits) {
        *link = r->next;
        if (s->tail == r) s->tail = prev;
        r->next = 0;
        r->state = 1;
        s->inflight_bytes += r->footprint;
        s->queued--;
        batch[n++] = r;
        continue;
      }
      if (r->table == table && !fits) break; /* keep arrival order on a table */
      prev = r;
      link = &r->next;
    }
    if (s->inflight_bytes > s->inflight_peak_bytes) s->inflight_peak_bytes = s->inflight_bytes;
    if (n == 0) {
      /* everything queued waits on memory held by running batches */
      pthread_cond_wait(&s->done_cv, &s->lock);
      continue;
    }
    pthread_mutex_unlock(&s->lock);

    B32 ok;
    if (kind == QE_REQUEST_JOIN && n == 1) {
      QE_JoinQuery* j = batch[0]->join;
      memset(&qs[0], 0, sizeof(qs[0]));
      qs[0].ok = qe_aggregate_join(table, j->fact_key_col, j->dim, j->dim_key_col, j->dim_where, j->dim_where_count,
                                   j->group_cols, j->group_count, j->funcs, j->arg_cols, j->expr_count, j->cap,
                                   j->out_group_count, j->out_keys, j->out_cells, 0);
    } else if (kind == QE_REQUEST_JOIN) {
      for (U32 i = 0; i < n; i++) jq[i] = *batch[i]->join;
      QE_JoinBatchTrace jt;
      const B32 ran = qe_aggregate_join_batch(table, jq, n, &jt);
      for (U32 i = 0; i < n; i++) {
        memset(&qs[i], 0, sizeof(qs[i]));
        qs[i].ok = ran && jq[i].ok;
      }
    } else if (kind == QE_REQUEST_SCAN) {
      QE_ScanQuery* q = batch[0]->scan;
      memset(&qs[0], 0, sizeof(qs[0]));
      qs[0].ok = qe_scan_filter(table, q->where, q->where_count, q->count_only, q->count_only ? 0 : q->out_ids,
                                q->count_only ? 0 : q->cap, q->out_count, 0);
    } else {
      for (U32 i = 0; i < n; i++) qs[i] = *batch[i]->q;
    }
    if (kind != QE_REQUEST_AGGREGATE) {
      /* ran above */
    } else if (n == 1 && s->max_batch == 1) {
      /* one query per launch: the plain operator */
      QE_BatchQuery* q = &qs[0];
      q->ok = qe_aggregate(table, q->where, q->where_count, q->group_cols, q->group_count, q->funcs, q->arg_cols,
                           q->expr_count, q->cap,
                           q->out_group_count, q->out_keys, q->out_cells, 0);
      q->fused = 0;
      ok = q->ok;
    } else {
      ok = qe_aggregate_batch(table, qs, n, 0);
    }
    (void)ok;
    U64 now = qe_now_ns();

    pthread_mutex_lock(&s->lock);
    for (U32 i = 0; i < n; i++) {
      QE_Request* r = batch[i];
      if (kind == QE_REQUEST_JOIN) {
        r->join->ok = qs[i].ok;
        s->joins += qs[i].ok ? 1 : 0;
      } else if (kind == QE_REQUEST_SCAN) {
        r->scan->ok = qs[i].ok;
        s->scans += qs[i].ok ? 1 : 0;
      } else {
        r->q->ok = qs[i].ok;
        r->q->fused = qs[i].fused;
      }
      r->t_done = now;
      r->state = 2;
      s->inflight_bytes -= r->footprint;
      if (qs[i].ok) {
        s->completed++;
        if (s->lat_count == s->lat_cap) {
          U64 cap = s->lat_cap ? s->lat_cap * 2 : 4096;
          U64* grown = (U64*)realloc(s->lat, cap * sizeof(U64));
          if (grown) {
            s->lat = grown;
            s->lat_cap = cap;
          }
        }
        if (s->lat_count < s->lat_cap) s->lat[s->lat_count++] = now - r->t_submit;
      } else {
        s->failed++;
      }
    }
    if (kind == QE_REQUEST_AGGREGATE) {
      s->batches++;
      s->batched_queries += n;
    }
    s->t_last_done = now;
    pthread_cond_broadcast(&s->done_cv);
    pthread_cond_broadcast(&s->work_cv);
  }
  pthread_mutex_unlock(&s->lock);
  free(batch);
  free(qs);
  free(jq);
  return 0;
}

QE_Server* qe_server_start(const QE_ServerConfig* config)
{
  QE_Server* s = (QE_Server*)calloc(1, sizeof(QE_Server));
  if (!s) {
    qe_set_error("qe_server_start: out of memory");
    return 0;
  }
  s->workers = config && config->workers ? config->workers : 2;
  if (s->workers > QE_SERVE_MAX_WORKERS) s->workers = QE_SERVE_MAX_WORKERS;
  s->max_batch = config && config->max_batch ? config->max_batch : 64;
  s->hbm_budget = config && config->hbm_budget ? config->hbm_budget : (1ull << 30);
  pthread_mutex_init(&s->lock, 0);
  pthread_cond_init(&s->work_cv, 0);
  pthread_cond_init(&s->done_cv, 0);
  for (U32 i = 0; i < s->workers; i++) {
    if (pthread_create(&s->threads[i], 0, qe_serve_worker, s) != 0) {
      s->workers = i;
      qe_server_stop(s);
      qe_set_error("qe_server_start: cannot start worker threads");
      return 0;
    }
  }
  return s;
}

static U64 qe_server_enqueue(QE_Server* s, QE_Request* r);

U64 qe_server_submit(QE_Server* s, QE_Table* table, QE_BatchQuery* query)
{
  if (!s || !table || !query || !query->funcs || !query->arg_cols || !query->out_group_count || !query->out_cells) {
    qe_set_error("qe_server_submit: null argument");
    return 0;
  }
  if (!qe_require_device(table, "qe_server_submit")) return 0;
  QE_Request* r = (QE_Request*)calloc(1, sizeof(QE_Request));
  if (!r) {
    qe_set_error("qe_server_submit: out of memory");
    return 0;
  }
  r->table = table;
  r->kind = QE_REQUEST_AGGREGATE;
  r->q = query;
  r->footprint = qe_serve_footprint(table, query);
  query->ok = 0;
  return qe_server_enqueue(s, r);
}

U64 qe_server_submit_join(QE_Server* s, QE_Table* fact, QE_JoinQuery* query)
{
  if (!s || !fact || !query || !query->dim || !query->funcs || !query->arg_cols || !query->out_group_count ||
      !query->out_cells || query->group_count > QE_AGG_MAX_GROUP_COLS) {
    qe_set_error("qe_server_submit_join: bad argument");
    return 0;
  }
  if (!qe_require_device(fact, "qe_server_submit_join") || !qe_require_device(query->dim, "qe_server_submit_join"))
    return 0;
  QE_Request* r = (QE_Request*)calloc(1, sizeof(QE_Request));
  if (!r) {
    qe_set_error("qe_server_submit_join: out of memory");
    return 0;
  }
  r->table = fact;
  r->kind = QE_REQUEST_JOIN;
  r->join = query;
  r->footprint = qe_serve_footprint_join(fact, query);
  query->ok = 0;
  return qe_server_enqueue(s, r);
}

U64 qe_server_submit_scan(QE_Server* s, QE_Table* table, QE_ScanQuery* query)
{
  if (!s || !table || !query || !query->out_count || (!query->count_only && !query->out_ids)) {
    qe_set_error("qe_server_submit_scan: null argument");
    return 0;
  }
  if (!qe_require_device(table, "qe_server_submit_scan")) return 0;
  QE_Request* r = (QE_Request*)calloc(1, sizeof(QE_Request));
  if (!r) {
    qe_set_error("qe_server_submit_scan: out of memory");
    return 0;
  }
  r->table = table;
  r->kind = QE_REQUEST_SCAN;
  r->scan = query;
  r->footprint = qe_serve_footprint_scan(table, query);
  query->ok = 0;
  return qe_server_enqueue(s, r);
}

static U64 qe_server_enqueue(QE_Server* s, QE_Request* r)
{
  r->t_submit = qe_now_ns();
  pthread_mutex_lock(&s->lock);
  if (s->stop) {
    pthread_mutex_unlock(&s->lock);
    free(r);
    qe_set_error("qe_server_submit: server stopped");
    return 0;
  }
  if (s->tail) s->tail->next = r;
  else s->head = r;
  s->tail = r;
  s->queued++;
  if (s->queued > s->queue_peak) s->queue_peak = s->queued;
  if (!s->submitted) s->t_first_submit = r->t_submit;
  s->submitted++;
  pthread_cond_signal(&s->work_cv);
  pthread_mutex_unlock(&s->lock);
  return (U64)(uintptr_t)r;
}

B32 qe_server_wait(QE_Server* s, U64 ticket, U64 timeout_ns)
{
  QE_Request* r = (QE_Request*)(uintptr_t)ticket;
  if (!s || !r) return 0;
  struct timespec until;
  clock_gettime(CLOCK_REALTIME, &until);
  U64 ns = (U64)until.tv_nsec + timeout_ns;
  until.tv_sec += (time_t)(ns / 1000000000ull);
  until.tv_nsec = (long)(ns % 1000000000ull);
  pthread_mutex_lock(&s->lock);
  while (r->state != 2) {
    if (pthread_cond_timedwait(&s->done_cv, &s->lock, &until) != 0 && r->state != 2) {
      pthread_mutex_unlock(&s->lock);
      return 0;
    }
  }
  pthread_mutex_unlock(&s->lock);
  free(r);
  return 1;
}

void qe_server_stats(QE_Server* s, QE_ServerStats* out)
{
  if (!s || !out) return;
  memset(out, 0, sizeof(*out));
  pthread_mutex_lock(&s->lock);
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->failed = s->failed;
  out->batches = s->batches;
  out->batched_queries = s->batched_queries;
  out->elapsed_ns = s->t_last_done > s->t_first_submit ? s->t_last_done - s->t_first_submit : 0;
  out->inflight_peak_bytes = s->inflight_peak_bytes;
  out->queue_peak = s->queue_peak;
  out->joins = s->joins;
  out->scans = s->scans;
  U64 n = s->lat_count;
  U64* c = n ? (U64*)malloc(n * sizeof(U64)) : 0;
  if (c) memcpy(c, s->lat, n * sizeof(U64));
  pthread_mutex_unlock(&s->lock);
  if (c) {
    qsort(c, n, sizeof(U64), qe_u64_cmp);
    U64 sum = 0;
    for (U64 i = 0; i < n; i++) sum += c[i];
    out->mean_ns = sum / n;
    out->p50_ns = c[(n - 1) / 2];
    out->p99_ns = c[(U64)((F64)(n - 1) * 0.99)];
    out->max_ns = c[n - 1];
    free(c);
  }
}

void qe_server_stats_reset(QE_Server* s)
{
  if (!s) return;
  pthread_mutex_lock(&s->lock);
  s->submitted = s->completed = s->failed = s->batches = s->batched_queries = s->joins = s->scans = 0;
  s->t_first_submit = s->t_last_done = 0;
  s->inflight_peak_bytes = s->inflight_bytes;
  s->queue_peak = s->queued;
  s->lat_count = 0;
  pthread_mutex_unlock(&s->lock);
}

void qe_server_stop(QE_Server* s)
{
  if (!s) return;
  pthread_mutex_lock(&s->lock);
  s->stop = 1;
  pthread_cond_broadcast(&s->work_cv);
  pthread_cond_broadcast(&s->done_cv);
  pthread_mutex_unlock(&s->lock);
  for (U32 i = 0; i < s->workers; i++) pthread_join(s->threads[i], 0);
  /* requests never run (none when the workers drained the queue) */
  QE_Request* r = s->head;
  while (r) {
    QE_Request* n = r->next;
    free(r);
    r = n;
  }
  free(s->lat);
  pthread_cond_destroy(&s->work_cv);
  pthread_cond_destroy(&s->done_cv);
  pthread_mutex_destroy(&s->lock);
  free(s);
}

/* ---------------- load driver ---------------- */
/* Poisson arrivals at `rate_qps` for `duration_ns`, each request a copy of
 * pool[k % pool_count] (k = arrival number) with output buffers of its own;
 * a collector thread waits for the tickets in arrival order.  Stats are
 * reset before the first arrival. */
typedef struct QE_DriveSlot {
  U64 ticket;
  U32 kind;
  QE_BatchQuery q;
  QE_JoinQuery j;
  QE_ScanQuery s;
  U64 group_count;
  void* out_a; /* keys or ids */
  void* out_b; /* cells */
  struct QE_DriveSlot* next;
} QE_DriveSlot;

static void qe_drive_slot_free(QE_DriveSlot* slot)
{
  free(slot->out_a);
  free(slot->out_b);
  free(slot);
}

/* a copy of the pool entry with output buffers of its own, submitted */
static QE_DriveSlot* qe_drive_submit(QE_Server* s, const QE_ServeRequest* p)
{
  QE_DriveSlot* slot = (QE_DriveSlot*)calloc(1, sizeof(QE_DriveSlot));
  if (!slot) return 0;
  slot->kind = p->kind;
  if (p->kind == QE_REQUEST_SCAN) {
    slot->s = *(const QE_ScanQuery*)p->query;
    const U64 cap = slot->s.count_only ? 0 : slot->s.cap;
    slot->out_a = malloc((cap ? cap : 1) * sizeof(U64));
    slot->s.out_ids = (U64*)slot->out_a;
    slot->s.out_count = &slot->group_count;
    if (slot->out_a) slot->ticket = qe_server_submit_scan(s, p->table, &slot->s);
  } else {
    const U32 groups = p->kind == QE_REQUEST_JOIN ? ((const QE_JoinQuery*)p->query)->group_count
                                                  : ((const QE_BatchQuery*)p->query)->group_count;
    const U32 exprs = p->kind == QE_REQUEST_JOIN ? ((const QE_JoinQuery*)p->query)->expr_count
                                                 : ((const QE_BatchQuery*)p->query)->expr_count;
    U64 cap = p->kind == QE_REQUEST_JOIN ? ((const QE_JoinQuery*)p->query)->cap
                                         : ((const QE_BatchQuery*)p->query)->cap;
    if (!cap) cap = 1;
    slot->out_a = malloc(cap * (groups ? groups : 1) * sizeof(F64));
    slot->out_b = malloc(cap * (exprs ? exprs : 1) * sizeof(U64));
    if (slot->out_a && slot->out_b) {
      if (p->kind == QE_REQUEST_JOIN) {
        slot->j = *(const QE_JoinQuery*)p->query;
        slot->j.out_group_count = &slot->group_count;
        slot->j.out_keys = (F64*)slot->out_a;
        slot->j.out_cells = (U64*)slot->out_b;
        slot->ticket = qe_server_submit_join(s, p->table, &slot->j);
      } else {
        slot->q = *(const QE_BatchQuery*)p->query;
        slot->q.out_group_count = &slot->group_count;
        slot->q.out_keys = (F64*)slot->out_a;
        slot->q.out_cells = (U64*)slot->out_b;
        slot->ticket = qe_server_submit(s, p->table, &slot->q);
      }
    }
  }
  if (!slot->ticket) {
    qe_drive_slot_free(slot);
    return 0;
  }
  return slot;
}

typedef struct QE_Drive {
  QE_Server* server;
  pthread_mutex_t lock;
  pthread_cond_t cv;
  QE_DriveSlot* head;
  QE_DriveSlot* tail;
  B32 done;
  U64 lost;
} QE_Drive;

static void* qe_drive_collect(void* arg)
{
  QE_Drive* d = (QE_Drive*)arg;
  for (;;) {
    pthread_mutex_lock(&d->lock);
    while (!d->head && !d->done) pthread_cond_wait(&d->cv, &d->lock);
    QE_DriveSlot* slot = d->head;
    if (!slot) {
      pthread_mutex_unlock(&d->lock);
      break;
    }
    d->head = slot->next;
    if (!d->head) d->tail = 0;
    pthread_mutex_unlock(&d->lock);
    if (!qe_server_wait(d->server, slot->ticket, 600ull * 1000000000ull)) {
      d->lost++; /* not released: its buffers stay allocated */
      continue;
    }
    qe_drive_slot_free(slot);
  }
  return 0;
}

static U64 qe_drive_rand(U64* state)
{
  U64 z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static B32 qe_server_drive_impl(QE_Server* s, const QE_ServeRequest* pool, U32 pool_count, F64 rate_qps,
                                 U64 duration_ns, U64 seed, QE_ServerStats* out);

/* the aggregate pools of qe_server_drive / _mix as request pools */
static B32 qe_server_drive_aggregates(QE_Server* s, QE_Table* table, QE_Table* const* tables,
                                      const QE_BatchQuery* pool, U32 pool_count, F64 rate_qps, U64 duration_ns,
                                      U64 seed, QE_ServerStats* out)
{
  if (!pool || !pool_count) {
    qe_set_error("qe_server_drive: bad argument");
    return 0;
  }
  QE_ServeRequest* req = (QE_ServeRequest*)calloc(pool_count, sizeof(QE_ServeRequest));
  if (!req) {
    qe_set_error("qe_server_drive: out of memory");
    return 0;
  }
  for (U32 i = 0; i < pool_count; i++) {
    req[i].kind = QE_REQUEST_AGGREGATE;
    req[i].table = tables ? tables[i] : table;
    req[i].query = &pool[i];
  }
  const B32 ok = qe_server_drive_impl(s, req, pool_count, rate_qps, duration_ns, seed, out);
  free(req);
  return ok;
}

B32 qe_server_drive(QE_Server* s, QE_Table* table, const QE_BatchQuery* pool, U32 pool_count, F64 rate_qps,
                    U64 duration_ns, U64 seed, QE_ServerStats* out)
{
  if (!table) {
    qe_set_error("qe_server_drive: bad argument");
    return 0;
  }
  return qe_server_drive_aggregates(s, table, 0, pool, pool_count, rate_qps, duration_ns, seed, out);
}

/* the same load over a mix of tables: request k runs pool[k % pool_count]
 * on tables[k % pool_count] */
B32 qe_server_drive_mix(QE_Server* s, QE_Table* const* tables, const QE_BatchQuery* pool, U32 pool_count,
                        F64 rate_qps, U64 duration_ns, U64 seed, QE_ServerStats* out)
{
  if (!tables) {
    qe_set_error("qe_server_drive_mix: bad argument");
    return 0;
  }
  for (U32 i = 0; i < pool_count; i++) {
    if (!tables[i]) {
      qe_set_error("qe_server_drive_mix: no table for query %u", i);
      return 0;
    }
  }
  return qe_server_drive_aggregates(s, 0, tables, pool, pool_count, rate_qps, duration_ns, seed, out);
}

B32 qe_server_drive_requests(QE_Server* s, const QE_ServeRequest* pool, U32 pool_count, F64 rate_qps,
                             U64 duration_ns, U64 seed, QE_ServerStats* out)
{
  if (!pool) {
    qe_set_error("qe_server_drive_requests: bad argument");
    return 0;
  }
  for (U32 i = 0; i < pool_count; i++) {
    if (!pool[i].table || !pool[i].query || pool[i].kind > QE_REQUEST_SCAN) {
      qe_set_error("qe_server_drive_requests: bad request %u", i);
      return 0;
    }
  }
  return qe_server_drive_impl(s, pool, pool_count, rate_qps, duration_ns, seed, out);
}

static B32 qe_server_drive_impl(QE_Server* s, const QE_ServeRequest* pool, U32 pool_count, F64 rate_qps,
                                 U64 duration_ns, U64 seed, QE_ServerStats* out)
{
  if (!s || !pool || !pool_count || rate_qps <= 0.0) {
    qe_set_error("qe_server_drive: bad argument");
    return 0;
  }
  QE_Drive d;
  memset(&d, 0, sizeof(d));
  d.server = s;
  pthread_mutex_init(&d.lock, 0);
  pthread_cond_init(&d.cv, 0);
  pthread_t collector;
  if (pthread_create(&collector, 0, qe_drive_collect, &d) != 0) {
    qe_set_error("qe_server_drive: cannot start the collector");
    return 0;
  }
  qe_server_stats_reset(s);
  U64 rng = seed;
  U64 t0 = qe_now_ns();
  F64 next = 0.0;
  B32 ok = 1;
  for (U64 k = 0;; k++) {
    /* exponential gap, from a uniform in (0, 1] */
    F64 u = ((F64)(qe_drive_rand(&rng) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    next += -log(u) / rate_qps * 1e9;
    if (next >= (F64)duration_ns) break;
    U64 at = t0 + (U64)next;
    for (;;) {
      U64 now = qe_now_ns();
      if (now >= at) break;
      if (at - now > 200000) {
        struct timespec ts = {0, (long)(at - now - 100000)};
        nanosleep(&ts, 0);
      }
    }
    QE_DriveSlot* slot = qe_drive_submit(s, &pool[k % pool_count]);
    if (!slot) {
      ok = 0;
      break;
    }
    pthread_mutex_lock(&d.lock);
    if (d.tail) d.tail->next = slot;
    else d.head = slot;
    d.tail = slot;
    pthread_cond_signal(&d.cv);
    pthread_mutex_unlock(&d.lock);
  }
  pthread_mutex_lock(&d.lock);
  d.done = 1;
  pthread_cond_signal(&d.cv);
  pthread_mutex_unlock(&d.lock);
  pthread_join(collector, 0);
  pthread_cond_destroy(&d.cv);
  pthread_mutex_destroy(&d.lock);
  if (out) qe_server_stats(s, out);
  if (d.lost) {
    qe_set_error("qe_server_drive: %llu requests did not finish", (unsigned long long)d.lost);
    ok = 0;
  }
  return ok;
}
