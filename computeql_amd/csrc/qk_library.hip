/*
 * qk_library.hip - hand-written gfx950 kernels behind gpu_hip_kernel_alloc.
 *
 * Every kernel takes one QK_Args by value (16 buffer bindings with byte
 * sizes + 16 push words) and is launched with 256-thread workgroups (four
 * wave64s).  The kernels are grid-size agnostic: the host picks the grid
 * (normally a few blocks per CU), and each block walks one contiguous range.
 *
 * Counterparts in the reference (GLSL, src/gpu/vulkan/shaders/):
 *   fill_u32            vkCmdFillBuffer of batch_buffer_fill
 *   bandwidth_read      bandwidth_read.comp
 *   scan_simple         scan_simple.comp
 *   scan_filter         scan_filter.comp (bytecode interpreter)
 *   aggregate_global    aggregate_global_native.comp / _f32.comp, fused with
 *                       an optional numeric-AND WHERE (fused global aggregate)
 *   aggregate_dense     aggregate_dense_native.comp, with an optional WHERE
 *   join_build          host table build of qe_join_probe_stream_core
 *                       (query_exec.c:8149-8191), done on the device
 *   join_probe          join_probe_native.comp
 *   gen_* / zonemap_build  data generators and zone maps (cql.c:3818)
 */
#include <hip/hip_runtime.h>

#include "../../include/cqlhip.h"
#include "qk_device_common.h"
#include "qk_library.h"

#define LB __launch_bounds__(QK_WG)

/* ------------------------------------------------------------------ */
/* fill_u32: binding 0 = buffer; push 0 = words, push 1 = value         */
/* ------------------------------------------------------------------ */
extern "C" __global__ void LB k_fill_u32(QK_Args a)
{
  qk_u32* p = qk_wptr<qk_u32>(a, 0);
  qk_u64 n = qk_push(a, 0);
  qk_u32 v = (qk_u32)qk_push(a, 1);
  qk_u64 stride = (qk_u64)gridDim.x * QK_WG;
  for (qk_u64 i = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; i < (n + 3) / 4; i += stride) {
    if (i * 4 + 3 < n) {
      ((qk_u32x4*)p)[i] = qk_u32x4{v, v, v, v};
    } else {
      for (qk_u64 k = i * 4; k < n; k++) p[k] = v;
    }
  }
}

/* ------------------------------------------------------------------ */
/* bandwidth_read: binding 0 = data, binding 1 = sink (u64);            */
/* push 0 = 16-byte units to read                                       */
/* ------------------------------------------------------------------ */
/* bandwidth probes: contiguous block ranges (as the scans walk them) in
 * steps of WG x QK_BW_UNROLL 16-byte units, every load of a step in flight
 * together.  bandwidth_copy: binding 0 source, 1 destination, push 0 units;
 * every unit read and written back once. */
#define QK_BW_UNROLL 8

__device__ __forceinline__ void qk_bw_range(qk_u64 n, qk_u64& b0, qk_u64& b1)
{
  const qk_u64 step = (qk_u64)QK_WG * QK_BW_UNROLL;
  qk_u64 per = (n + gridDim.x - 1) / gridDim.x;
  per = (per + step - 1) / step * step;
  b0 = per * blockIdx.x;
  if (b0 > n) b0 = n;
  b1 = b0 + per < n ? b0 + per : n;
}

extern "C" __global__ void LB k_bandwidth_copy(QK_Args a)
{
  const qk_u32x4* p = qk_ptr<qk_u32x4>(a, 0);
  qk_u32x4* o = qk_wptr<qk_u32x4>(a, 1);
  qk_u64 b0, b1;
  qk_bw_range(qk_push(a, 0), b0, b1);
  for (qk_u64 s = b0; s < b1; s += (qk_u64)QK_WG * QK_BW_UNROLL) {
    qk_u32x4 x[QK_BW_UNROLL];
#pragma unroll
    for (int j = 0; j < QK_BW_UNROLL; j++) {
      qk_u64 i = s + (qk_u64)j * QK_WG + threadIdx.x;
      x[j] = i < b1 ? qk_ld(p + i) : qk_u32x4{0, 0, 0, 0};
    }
#pragma unroll
    for (int j = 0; j < QK_BW_UNROLL; j++) {
      qk_u64 i = s + (qk_u64)j * QK_WG + threadIdx.x;
      if (i < b1) o[i] = x[j];
    }
  }
}

/* bandwidth_read: binding 0 the buffer, 1 a sink word; push 0 units */
extern "C" __global__ void LB k_bandwidth_read(QK_Args a)
{
  const qk_u32x4* p = qk_ptr<qk_u32x4>(a, 0);
  qk_u64 b0, b1;
  qk_bw_range(qk_push(a, 0), b0, b1);
  qk_u32x4 acc = {0, 0, 0, 0};
  for (qk_u64 s = b0; s < b1; s += (qk_u64)QK_WG * QK_BW_UNROLL) {
    qk_u32x4 x[QK_BW_UNROLL];
#pragma unroll
    for (int j = 0; j < QK_BW_UNROLL; j++) {
      qk_u64 i = s + (qk_u64)j * QK_WG + threadIdx.x;
      x[j] = i < b1 ? qk_ld(p + i) : qk_u32x4{0, 0, 0, 0};
    }
#pragma unroll
    for (int j = 0; j < QK_BW_UNROLL; j++) acc ^= x[j];
  }
  qk_u32 r = acc.x ^ acc.y ^ acc.z ^ acc.w;
  if (r == 0x9e3779b9u) qk_wptr<qk_u64>(a, 1)[0] = r; /* keeps the loads live */
}

/* ------------------------------------------------------------------ */
/* aggregate_global: COUNT and SUM/MIN/MAX of up to 4 columns with an   */
/* optional numeric-AND WHERE, one pass.                                */
/*  binding 0 leaves (6 words each), 2 partials (grid x 13 u64),       */
/*  5 skip bitmap, 6.. column slots                                     */
/*  push 0 rows, 3 skip shift, 4 skip on, 5 first row, 7 leaf count     */
/*  words[12] = arg count | arg slots << 8 (4 bits each)                */
/*  words[13] = arg types (8 bits each)                                 */
/* partials, per block: [0] count, [1+3c] sum, [2+3c] min key, [3+3c]  */
/* max key.  Sums: exact 64-bit integers for U32/I32 (I32 as two's      */
/* complement), f64 bits otherwise.  Min/max keys are order-preserving  */
/* u64 encodings (u32: value; i32/i64: value ^ sign; f64: IEEE flip).   */
/* aggregate_global_fold then folds the partials into the result in a   */
/* fixed order, so the f64 result depends only on the grid size.        */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ qk_u64 qk_f64_key(double d)
{
  qk_u64 b = (qk_u64)__double_as_longlong(d);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}

__device__ __forceinline__ bool qk_type_is_int32(qk_u32 t) { return t == QK_T_U32 || t == QK_T_I32; }

struct AggAcc {
  qk_u64 sum, mn, mx;
};

__device__ __forceinline__ void agg_fold_quad(const QK_Args& a, qk_u32 slot, qk_u32 type, qk_u64 q, qk_u32 m,
                                              AggAcc& acc)
{
  qk_u64 base = a.ptr[QK_B_COL0 + slot];
  if (type == QK_T_U32) {
    qk_u32x4 x = qk_ld(qk_gptr<const qk_u32x4>(base) + q);
    if (m == 0xfu) {
      acc.sum += (qk_u64)x.x + (qk_u64)x.y + (qk_u64)x.z + (qk_u64)x.w;
      qk_u32 lo = min(min(x.x, x.y), min(x.z, x.w));
      qk_u32 hi = max(max(x.x, x.y), max(x.z, x.w));
      acc.mn = min(acc.mn, (qk_u64)lo);
      acc.mx = max(acc.mx, (qk_u64)hi);
    } else {
      qk_u32 w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (m & (1u << k)) {
          acc.sum += (qk_u64)w[k];
          acc.mn = min(acc.mn, (qk_u64)w[k]);
          acc.mx = max(acc.mx, (qk_u64)w[k]);
        }
      }
    }
  } else if (type == QK_T_I32) {
    qk_u32x4 x = qk_ld(qk_gptr<const qk_u32x4>(base) + q);
    qk_u32 w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (m & (1u << k)) {
        acc.sum += (qk_u64)(qk_s64)(qk_s32)w[k];
        qk_u64 key = (qk_u64)(w[k] ^ 0x80000000u);
        acc.mn = min(acc.mn, key);
        acc.mx = max(acc.mx, key);
      }
    }
  } else {
    double v[4];
    qk_load4_f64(a, slot, type, q, v);
    double s = __longlong_as_double((qk_s64)acc.sum);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (m & (1u << k)) {
        s += v[k];
        qk_u64 key;
        if (type == QK_T_U64) {
          qk_u64x2 x = qk_gptr<const qk_u64x2>(base)[2 * q + (k >> 1)];
          key = (k & 1) ? x.y : x.x;
        } else if (type == QK_T_I64) {
          qk_u64x2 x = qk_gptr<const qk_u64x2>(base)[2 * q + (k >> 1)];
          key = ((k & 1) ? x.y : x.x) ^ 0x8000000000000000ULL;
        } else {
          key = qk_f64_key(v[k]);
        }
        acc.mn = min(acc.mn, key);
        acc.mx = max(acc.mx, key);
      }
    }
    acc.sum = (qk_u64)__double_as_longlong(s);
  }
}

__device__ __forceinline__ qk_u64 agg_combine_sum(qk_u32 type, qk_u64 x, qk_u64 y)
{
  if (qk_type_is_int32(type)) return x + y;
  return (qk_u64)__double_as_longlong(__longlong_as_double((qk_s64)x) + __longlong_as_double((qk_s64)y));
}

template <bool TERMS>
__device__ __forceinline__ void qk_aggregate_global(const QK_Args& a)
{
  __shared__ qk_u64 s_part[QK_WG / 64][13];
  const qk_u64 rows = qk_push(a, QK_P_ROWS);
  const qk_u64 r0 = qk_push(a, QK_P_ROW_BEGIN);
  const qk_u64 r1 = r0 + rows;
  const qk_u32 leaf_count = (qk_u32)qk_push(a, QK_P_LEAVES);
  const qk_u32 nargs = a.words[QK_W_ARGS] & 0xffu;
  const qk_u32 slots = a.words[QK_W_ARGS] >> 8;
  const qk_u32 types = a.words[QK_W_TYPES];
  const QkLeafSet L = qk_load_leaves<TERMS>(qk_ptr<qk_u32>(a, QK_B_LEAVES), leaf_count);

  qk_u64 count = 0;
  AggAcc acc[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    acc[c].sum = 0;
    acc[c].mn = ~0ULL;
    acc[c].mx = 0;
  }
  qk_u64 b0, b1;
  qk_block_range(r0 / 4, (r1 + 3) / 4, b0, b1);
  for (qk_u64 s = b0; s < b1; s += (qk_u64)QK_WG * QK_UNROLL) {
    qk_u64 q[QK_UNROLL];
    qk_u32 m[QK_UNROLL];
    qk_step_quads(a, s, b1, r0, r1, q, m);
    if (L.n) qk_leaves_step<TERMS>(a, L, q, m);
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) count += (qk_u64)__popc(m[j]);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if ((qk_u32)c >= nargs) continue;
      /* type byte: column type, encoding << 4 (encoded args are codes, U32) */
      qk_u32 slot = (slots >> (4 * c)) & 0xfu, type = (types >> (8 * c)) & 0xfu, enc = (types >> (8 * c + 4)) & 3u;
      if (type == QK_T_U32) {
        /* the hot case: the U loads of the step in flight together */
        qk_u32x4 x[QK_UNROLL];
        if (enc == 0) {
          const qk_u32x4* col = qk_ptr<qk_u32x4>(a, QK_B_COL0 + slot);
#pragma unroll
          for (int j = 0; j < QK_UNROLL; j++) x[j] = qk_ld(col + q[j]);
        } else {
#pragma unroll
          for (int j = 0; j < QK_UNROLL; j++) x[j] = qk_quad_codes(a.ptr[QK_B_COL0 + slot], enc, q[j]);
        }
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) {
          qk_u32 mm = m[j];
          qk_u32 w0 = (mm & 1u) ? x[j].x : 0u, w1 = (mm & 2u) ? x[j].y : 0u;
          qk_u32 w2 = (mm & 4u) ? x[j].z : 0u, w3 = (mm & 8u) ? x[j].w : 0u;
          acc[c].sum += (qk_u64)w0 + (qk_u64)w1 + (qk_u64)w2 + (qk_u64)w3;
          qk_u32 lo = min(min((mm & 1u) ? x[j].x : 0xffffffffu, (mm & 2u) ? x[j].y : 0xffffffffu),
                          min((mm & 4u) ? x[j].z : 0xffffffffu, (mm & 8u) ? x[j].w : 0xffffffffu));
          qk_u32 hi = max(max(w0, w1), max(w2, w3));
          if (mm) {
            acc[c].mn = min(acc[c].mn, (qk_u64)lo);
            acc[c].mx = max(acc[c].mx, (qk_u64)hi);
          }
        }
      } else {
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++)
          if (m[j]) agg_fold_quad(a, slot, type, q[j], m[j], acc[c]);
      }
    }
  }

  /* wave reduce, then across the 4 waves in wave order */
  qk_u32 lane = __lane_id(), wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    count += __shfl_xor(count, o, 64);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      qk_u32 t = (types >> (8 * c)) & 0xfu;
      qk_u64 os = __shfl_xor(acc[c].sum, o, 64);
      acc[c].sum = agg_combine_sum(t, acc[c].sum, os);
      acc[c].mn = min(acc[c].mn, (qk_u64)__shfl_xor(acc[c].mn, o, 64));
      acc[c].mx = max(acc[c].mx, (qk_u64)__shfl_xor(acc[c].mx, o, 64));
    }
  }
  if (lane == 0) {
    s_part[wave][0] = count;
    for (int c = 0; c < 4; c++) {
      s_part[wave][1 + 3 * c] = acc[c].sum;
      s_part[wave][2 + 3 * c] = acc[c].mn;
      s_part[wave][3 + 3 * c] = acc[c].mx;
    }
  }
  __syncthreads();
  qk_u64* partials = qk_wptr<qk_u64>(a, 2);
  if (threadIdx.x < 13) {
    int f = threadIdx.x;
    qk_u64 v = s_part[0][f];
    for (int w = 1; w < QK_WG / 64; w++) {
      qk_u64 x = s_part[w][f];
      if (f == 0) v += x;
      else {
        int c = (f - 1) / 3, kind = (f - 1) % 3;
        qk_u32 t = (types >> (8 * c)) & 0xfu;
        v = kind == 0 ? agg_combine_sum(t, v, x) : (kind == 1 ? min(v, x) : max(v, x));
      }
    }
    partials[(qk_u64)blockIdx.x * 13 + f] = v;
  }
}

extern "C" __global__ void LB k_aggregate_global(QK_Args a) { qk_aggregate_global<false>(a); }
/* the same over a WHERE of several conjunctions */
extern "C" __global__ void LB k_aggregate_global_or(QK_Args a) { qk_aggregate_global<true>(a); }

/* aggregate_global_fold: one block folds aggregate_global's per-block
 * partials (binding 2, 13 words a block; push 0 = blocks) into binding 1's
 * words 0..12, in a fixed order: thread t walks blocks t, t+256, ... in
 * order, then the 256 partials meet in a fixed butterfly + wave order.  It
 * runs as its own launch after aggregate_global: the kernel boundary makes
 * the partials visible, where a last-block fold inside aggregate_global
 * would need an agent-scope release/acquire per block (an L2 write-back
 * and invalidate on this part, under the blocks still streaming).
 * Words ARGS, TYPES as aggregate_global. */
extern "C" __global__ void LB k_aggregate_global_fold(QK_Args a)
{
  __shared__ qk_u64 s_part[QK_WG / 64][13];
  const qk_u32 nblocks = (qk_u32)qk_push(a, 0);
  const qk_u32 nargs = a.words[QK_W_ARGS] & 0xffu;
  const qk_u32 types = a.words[QK_W_TYPES];
  const qk_u64* partials = qk_ptr<qk_u64>(a, 2);
  qk_u64* result = qk_wptr<qk_u64>(a, 1);
  const qk_u32 lane = __lane_id(), wave = threadIdx.x >> 6;
  qk_u64 v[13];
  v[0] = 0;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    v[1 + 3 * c] = 0;
    v[2 + 3 * c] = ~0ULL;
    v[3 + 3 * c] = 0;
  }
  for (qk_u32 b = threadIdx.x; b < nblocks; b += QK_WG) {
    const qk_u64* pb = &partials[(qk_u64)b * 13];
    v[0] += pb[0];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if ((qk_u32)c >= nargs) continue;
      qk_u32 t = (types >> (8 * c)) & 0xfu;
      v[1 + 3 * c] = agg_combine_sum(t, v[1 + 3 * c], pb[1 + 3 * c]);
      v[2 + 3 * c] = min(v[2 + 3 * c], pb[2 + 3 * c]);
      v[3 + 3 * c] = max(v[3 + 3 * c], pb[3 + 3 * c]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v[0] += __shfl_xor(v[0], o, 64);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      qk_u32 t = (types >> (8 * c)) & 0xfu;
      v[1 + 3 * c] = agg_combine_sum(t, v[1 + 3 * c], __shfl_xor(v[1 + 3 * c], o, 64));
      v[2 + 3 * c] = min(v[2 + 3 * c], (qk_u64)__shfl_xor(v[2 + 3 * c], o, 64));
      v[3 + 3 * c] = max(v[3 + 3 * c], (qk_u64)__shfl_xor(v[3 + 3 * c], o, 64));
    }
  }
  if (lane == 0)
    for (int f = 0; f < 13; f++) s_part[wave][f] = v[f];
  __syncthreads();
  if (threadIdx.x < 13) {
    int f = threadIdx.x;
    qk_u64 r = s_part[0][f];
    for (int w = 1; w < QK_WG / 64; w++) {
      qk_u64 x = s_part[w][f];
      if (f == 0) r += x;
      else {
        int c = (f - 1) / 3, kind = (f - 1) % 3;
        qk_u32 t = (types >> (8 * c)) & 0xfu;
        r = kind == 0 ? agg_combine_sum(t, r, x) : (kind == 1 ? min(r, x) : max(r, x));
      }
    }
    result[f] = r;
  }
}

/* ------------------------------------------------------------------ */
/* scan_simple: numeric-AND leaves, count / block counts / ordered ids  */
/* (scan bindings and push slots of query_exec.h:4-40)                  */
/* ------------------------------------------------------------------ */
template <bool TERMS>
struct SimpleStepPred {
  const QK_Args* a;
  QkLeafSet L;
  __device__ __forceinline__ void eval(const qk_u64* q, const qk_u32* valid, qk_u32* m) const
  {
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) m[j] = valid[j];
    qk_leaves_step<TERMS>(*a, L, q, m);
  }
};

extern "C" __global__ void LB k_scan_simple(QK_Args a)
{
  SimpleStepPred<false> p{&a, qk_load_leaves<false>(qk_ptr<qk_u32>(a, QK_B_LEAVES), (qk_u32)qk_push(a, QK_P_LEAVES))};
  qk_scan_driver_step(a, p);
}

/* the same over a WHERE of several conjunctions */
extern "C" __global__ void LB k_scan_simple_or(QK_Args a)
{
  SimpleStepPred<true> p{&a, qk_load_leaves<true>(qk_ptr<qk_u32>(a, QK_B_LEAVES), (qk_u32)qk_push(a, QK_P_LEAVES))};
  qk_scan_driver_step(a, p);
}

/* ------------------------------------------------------------------ */
/* scan_filter: the stack-bytecode interpreter (scan_filter.comp), kept  */
/* as the fallback for shapes the other kernels do not take.  Opcodes   */
/* of QE_Opcode (query_exec.h:58-82), numeric subset; stack of 8 f64.   */
/* binding 0 bytecode (u32 words), 1 numeric constants (f64).           */
/* ------------------------------------------------------------------ */
#define OP_PUSH_TRUE 0
#define OP_LOAD_NUM_COL 1
#define OP_PUSH_CONST 2
#define OP_CMP_EQ 3
#define OP_CMP_NE 4
#define OP_CMP_LT 5
#define OP_CMP_GT 6
#define OP_CMP_LE 7
#define OP_CMP_GE 8
#define OP_AND 9
#define OP_OR 10
#define OP_HALT 13
#define OP_PUSH_FALSE 14
#define QK_MAX_STACK 8

__device__ __forceinline__ double qk_load1_f64(const QK_Args& a, qk_u32 slot, qk_u32 type, qk_u64 row)
{
  qk_u64 base = a.ptr[QK_B_COL0 + slot];
  switch (type) {
    case QK_T_U32: return (double)qk_gptr<const qk_u32>(base)[row];
    case QK_T_I32: return (double)qk_gptr<const qk_s32>(base)[row];
    case QK_T_F32: return (double)qk_gptr<const float>(base)[row];
    case QK_T_F64: return qk_gptr<const double>(base)[row];
    case QK_T_I64: return (double)qk_gptr<const qk_s64>(base)[row];
    case QK_T_U64: return (double)qk_gptr<const qk_u64>(base)[row];
    default: return 0.0;
  }
}

struct InterpPred {
  const QK_Args* a;
  const qk_u32* code;
  const double* consts;
  qk_u32 code_words;
  qk_u32 const_count;
  __device__ __forceinline__ qk_u32 operator()(qk_u64 q, qk_u32 valid) const
  {
    qk_u32 m = 0;
    for (int k = 0; k < 4; k++) {
      if (!(valid & (1u << k))) continue;
      qk_u64 row = q * 4 + k;
      double st[QK_MAX_STACK];
      int sp = 0;
      qk_u32 ip = 0;
      bool ok = true;
      while (ip < code_words && ok) {
        qk_u32 op = code[ip++];
        if (op == OP_HALT) break;
        switch (op) {
          case OP_PUSH_TRUE: ok = sp < QK_MAX_STACK; if (ok) st[sp++] = 1.0; break;
          case OP_PUSH_FALSE: ok = sp < QK_MAX_STACK; if (ok) st[sp++] = 0.0; break;
          case OP_LOAD_NUM_COL: {
            ok = ip < code_words && sp < QK_MAX_STACK;
            if (!ok) break;
            qk_u32 operand = code[ip++];
            ok = (operand & 0xffu) < 10u;
            if (ok) st[sp++] = qk_load1_f64(*a, operand & 0xffu, operand >> 8, row);
          } break;
          case OP_PUSH_CONST: {
            ok = ip < code_words && sp < QK_MAX_STACK;
            if (!ok) break;
            qk_u32 ci = code[ip++];
            ok = ci < const_count;
            if (ok) st[sp++] = consts[ci];
          } break;
          default: {
            ok = sp >= 2;
            if (!ok) break;
            double r = st[--sp], l = st[--sp];
            bool v;
            switch (op) {
              case OP_CMP_EQ: v = l == r; break;
              case OP_CMP_NE: v = l != r; break;
              case OP_CMP_LT: v = l < r; break;
              case OP_CMP_GT: v = l > r; break;
              case OP_CMP_LE: v = l <= r; break;
              case OP_CMP_GE: v = l >= r; break;
              case OP_AND: v = (l != 0.0) && (r != 0.0); break;
              case OP_OR: v = (l != 0.0) || (r != 0.0); break;
              default: v = false; ok = false; break;
            }
            st[sp++] = v ? 1.0 : 0.0;
          } break;
        }
      }
      if (ok && sp >= 1 && st[sp - 1] != 0.0) m |= 1u << k;
    }
    return m;
  }
};

extern "C" __global__ void LB k_scan_filter(QK_Args a)
{
  InterpPred p{&a, qk_ptr<qk_u32>(a, QK_B_LEAVES), qk_ptr<double>(a, QK_B_CONSTS), (qk_u32)(a.bytes[QK_B_LEAVES] / 4),
               (qk_u32)(a.bytes[QK_B_CONSTS] / 8)};
  qk_scan_driver(a, p);
}

/* ------------------------------------------------------------------ */
/* aggregate_dense: GROUP BY one U32/I32 key of range <= LDS capacity    */
/* (aggregate_dense_native.comp), optional numeric-AND WHERE.           */
/*  binding 0 leaves, 1 global accumulators, 5 skip, 6.. columns        */
/*  push 0 rows, 3 skip shift, 4 skip on, 5 first row, 7 leaf count     */
/*  words[12] = arg count | arg slots << 8; words[13] = arg types       */
/*  words[14] = key slot | key type << 8 | flags << 16                  */
/*            flags: 1 sums, 2 min/max                                  */
/*  words[10] = key code minimum, words[11] = range                     */
/*  binding 2 (optional, several GROUP BY keys): u32 key count K <= 3,  */
/*  then per key three words: slot | type << 8 | encoding << 12, code   */
/*  minimum, range; a row's group is then the keys' digits (key 0 most  */
/*  significant), the product of the ranges being the table's range    */
/*  (query_exec.c:9462-9500 combines small-range keys the same way).    */
/* Global layout (range R, args E): count u64[R]; sum u64[E][R];        */
/* min u32[E][R]; max u32[E][R].  u32 keys of args are value for U32,   */
/* value ^ 0x80000000 for I32 (order preserving).  Block-private LDS    */
/* tables; counts u32, sums u64, then one global atomic per touched     */
/* group and field at the end.                                          */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ qk_u32 qk_code32(qk_u32 v, qk_u32 type) { return type == QK_T_I32 ? (v ^ 0x80000000u) : v; }

template <bool TERMS, bool WIDE>
__device__ __forceinline__ void qk_aggregate_dense(const QK_Args& a)
{
  extern __shared__ qk_u64 lds_u64[];
  const qk_u64 rows = qk_push(a, QK_P_ROWS);
  const qk_u64 r0 = qk_push(a, QK_P_ROW_BEGIN);
  const qk_u64 r1 = r0 + rows;
  const qk_u32 leaf_count = (qk_u32)qk_push(a, QK_P_LEAVES);
  const qk_u32 nargs = a.words[QK_W_ARGS] & 0xffu;
  const qk_u32 slots = a.words[QK_W_ARGS] >> 8;
  const qk_u32 types = a.words[QK_W_TYPES];
  const qk_u32 key_slot = a.words[QK_W_KEY] & 0xffu;
  /* type bytes: column type, encoding << 4 (encoded columns are codes, U32) */
  const qk_u32 key_type = (a.words[QK_W_KEY] >> 8) & 0xfu;
  const qk_u32 key_enc = (a.words[QK_W_KEY] >> 12) & 3u;
  const qk_u32 flags = a.words[QK_W_KEY] >> 16;
  const qk_u32 key_min = a.words[QK_W_KEY_MIN];
  const qk_u32 range = a.words[QK_W_RANGE];
  const QkLeafSet L = qk_load_leaves<TERMS>(qk_ptr<qk_u32>(a, QK_B_LEAVES), leaf_count);
  const qk_u32* kd = a.ptr[2] ? qk_ptr<qk_u32>(a, 2) : nullptr;
  const qk_u32 nkeys = kd ? kd[0] : 1u;

  /* LDS: sums u64[E][R] | counts u32[R] | min u32[E][R] | max u32[E][R];
   * WIDE (any argument wider than a 32-bit integer): sums u64[E][R] (exact
   * integers for U32/I32 arguments, f64 otherwise) | min u64[E][R] | max
   * u64[E][R] (order-preserving keys of the f64 value) | counts u32[R] */
  const qk_u32 sum_words = (flags & 1u) ? nargs * range : 0u;
  const qk_u32 mm_words = (flags & 2u) ? nargs * range : 0u;
  qk_u64* l_sum = lds_u64;
  qk_u64* l_min64 = l_sum + sum_words;
  qk_u64* l_max64 = l_min64 + mm_words;
  qk_u32* l_cnt = WIDE ? (qk_u32*)(l_max64 + mm_words) : (qk_u32*)(lds_u64 + sum_words);
  qk_u32* l_min = l_cnt + range;
  qk_u32* l_max = l_min + mm_words;
  for (qk_u32 i = threadIdx.x; i < sum_words; i += QK_WG) l_sum[i] = 0;
  for (qk_u32 i = threadIdx.x; i < range; i += QK_WG) l_cnt[i] = 0;
  for (qk_u32 i = threadIdx.x; i < mm_words; i += QK_WG) {
    if (WIDE) {
      l_min64[i] = ~0ULL;
      l_max64[i] = 0ULL;
    } else {
      l_min[i] = 0xffffffffu;
      l_max[i] = 0u;
    }
  }
  __syncthreads();

  const qk_u32x4* keys = qk_ptr<qk_u32x4>(a, QK_B_COL0 + key_slot);
  qk_u64 b0, b1;
  qk_block_range(r0 / 4, (r1 + 3) / 4, b0, b1);
  for (qk_u64 s = b0; s < b1; s += (qk_u64)QK_WG * QK_UNROLL) {
    qk_u64 q[QK_UNROLL];
    qk_u32 m[QK_UNROLL];
    qk_u32x4 kv[QK_UNROLL];
    qk_step_quads(a, s, b1, r0, r1, q, m);
    if (L.n) qk_leaves_step<TERMS>(a, L, q, m);
    qk_u32 g[QK_UNROLL][4];
    if (nkeys == 1) {
      if (key_enc == 0) {
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) kv[j] = qk_ld(keys + q[j]);
      } else {
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) kv[j] = qk_quad_codes(a.ptr[QK_B_COL0 + key_slot], key_enc, q[j]);
      }
#pragma unroll
      for (int j = 0; j < QK_UNROLL; j++) {
        g[j][0] = qk_code32(kv[j].x, key_type) - key_min;
        g[j][1] = qk_code32(kv[j].y, key_type) - key_min;
        g[j][2] = qk_code32(kv[j].z, key_type) - key_min;
        g[j][3] = qk_code32(kv[j].w, key_type) - key_min;
        /* a key outside [key_min, key_min + range) (stale zone map) is
         * dropped rather than written out of the table */
#pragma unroll
        for (int k = 0; k < 4; k++)
          if (g[j][k] >= range) m[j] &= ~(1u << k);
      }
    } else {
      /* several keys: digits, each checked against its own range */
#pragma unroll
      for (int j = 0; j < QK_UNROLL; j++) g[j][0] = g[j][1] = g[j][2] = g[j][3] = 0u;
      for (qk_u32 kk = 0; kk < nkeys; kk++) {
        const qk_u32 d = kd[1 + 3 * kk], kmin = kd[2 + 3 * kk], kr = kd[3 + 3 * kk];
        const qk_u32 kt = (d >> 8) & 0xfu, ke = (d >> 12) & 3u;
        const qk_u64 kp = a.ptr[QK_B_COL0 + (d & 0xffu)];
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) {
          const qk_u32x4 x = ke == 0 ? qk_ld(qk_gptr<const qk_u32x4>(kp) + q[j]) : qk_quad_codes(kp, ke, q[j]);
          const qk_u32 w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const qk_u32 dg = qk_code32(w[k], kt) - kmin;
            if (dg >= kr) m[j] &= ~(1u << k);
            g[j][k] = g[j][k] * kr + dg;
          }
        }
      }
#pragma unroll
      for (int j = 0; j < QK_UNROLL; j++)
#pragma unroll
        for (int k = 0; k < 4; k++)
          if (g[j][k] >= range) m[j] &= ~(1u << k);
    }
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++)
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (m[j] & (1u << k)) atomicAdd(&l_cnt[g[j][k]], 1u);
    for (qk_u32 c = 0; c < nargs; c++) {
      qk_u32 at = (types >> (8 * c)) & 0xfu, aenc = (types >> (8 * c + 4)) & 3u;
      const qk_u64 cp = a.ptr[QK_B_COL0 + ((slots >> (4 * c)) & 0xfu)];
      if (WIDE && !qk_type_is_int32(at)) {
        /* f32 / f64 / i64 / u64: the value as f64, summed as f64 */
        const qk_u32 aslot = (slots >> (4 * c)) & 0xfu;
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) {
          double v[4];
          qk_load4_f64(a, aslot, at, q[j], v);
#pragma unroll
          for (int k = 0; k < 4; k++) {
            if (!(m[j] & (1u << k))) continue;
            const qk_u64 i = (qk_u64)c * range + g[j][k];
            if (flags & 1u) atomicAdd((double*)&l_sum[i], v[k]);
            if (flags & 2u) {
              const qk_u64 key = qk_f64_key(v[k]);
              atomicMin(&l_min64[i], key);
              atomicMax(&l_max64[i], key);
            }
          }
        }
        continue;
      }
      qk_u32x4 x[QK_UNROLL];
      if (aenc == 0) {
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) x[j] = qk_ld(qk_gptr<const qk_u32x4>(cp) + q[j]);
      } else {
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) x[j] = qk_quad_codes(cp, aenc, q[j]);
      }
#pragma unroll
      for (int j = 0; j < QK_UNROLL; j++) {
        qk_u32 w[4] = {x[j].x, x[j].y, x[j].z, x[j].w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          if (!(m[j] & (1u << k))) continue;
          if (flags & 1u) {
            qk_u64 add = at == QK_T_I32 ? (qk_u64)(qk_s64)(qk_s32)w[k] : (qk_u64)w[k];
            atomicAdd(&l_sum[c * range + g[j][k]], add);
          }
          if (flags & 2u) {
            if (WIDE) {
              const qk_u64 key = qk_f64_key(at == QK_T_I32 ? (double)(qk_s32)w[k] : (double)w[k]);
              atomicMin(&l_min64[c * range + g[j][k]], key);
              atomicMax(&l_max64[c * range + g[j][k]], key);
            } else {
              qk_u32 code = qk_code32(w[k], at);
              atomicMin(&l_min[c * range + g[j][k]], code);
              atomicMax(&l_max[c * range + g[j][k]], code);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  /* global: count u64[R] | sum u64[E][R] | min, max u32[E][R] (WIDE: u64) */
  qk_u64* g_cnt = qk_wptr<qk_u64>(a, 1);
  qk_u64* g_sum = g_cnt + range;
  qk_u32* g_min = (qk_u32*)(g_sum + (qk_u64)nargs * range);
  qk_u32* g_max = g_min + (qk_u64)nargs * range;
  qk_u64* g_min64 = g_sum + (qk_u64)nargs * range;
  qk_u64* g_max64 = g_min64 + (qk_u64)nargs * range;
  for (qk_u32 i = threadIdx.x; i < range; i += QK_WG) {
    qk_u32 cnt = l_cnt[i];
    if (!cnt) continue;
    atomicAdd(&g_cnt[i], (qk_u64)cnt);
    for (qk_u32 c = 0; c < nargs; c++) {
      const qk_u64 gi = (qk_u64)c * range + i;
      if (flags & 1u) {
        if (WIDE && !qk_type_is_int32((types >> (8 * c)) & 0xfu))
          atomicAdd((double*)&g_sum[gi], __longlong_as_double((qk_s64)l_sum[c * range + i]));
        else
          atomicAdd(&g_sum[gi], l_sum[c * range + i]);
      }
      if (flags & 2u) {
        if (WIDE) {
          atomicMin(&g_min64[gi], l_min64[c * range + i]);
          atomicMax(&g_max64[gi], l_max64[c * range + i]);
        } else {
          atomicMin(&g_min[gi], l_min[c * range + i]);
          atomicMax(&g_max[gi], l_max[c * range + i]);
        }
      }
    }
  }
}

extern "C" __global__ void LB k_aggregate_dense(QK_Args a) { qk_aggregate_dense<false, false>(a); }
/* the same over a WHERE of several conjunctions */
extern "C" __global__ void LB k_aggregate_dense_or(QK_Args a) { qk_aggregate_dense<true, false>(a); }
/* arguments wider than 32-bit integers (f32, f64, i64, u64) */
extern "C" __global__ void LB k_aggregate_dense_wide(QK_Args a) { qk_aggregate_dense<false, true>(a); }
extern "C" __global__ void LB k_aggregate_dense_wide_or(QK_Args a) { qk_aggregate_dense<true, true>(a); }

/* ------------------------------------------------------------------ */
/* aggregate_dense_narrow: aggregate_dense for group ranges too large    */
/* for its LDS tables (10k groups x 20 B), with one U32 argument, in     */
/* 1024-thread blocks and at most 16 bytes a group:                      */
/*   cc u32[R]     count in bits 0-15, carries of the sum's low word in  */
/*                 bits 16-31                                             */
/*   lo u32[R]     the sum's low 32 bits (SUM)                            */
/*   mn, mx u32[R] (MM)                                                   */
/* The lane whose add takes a group's count half from 0x7fff to 0x8000,  */
/* or its carry half likewise, moves 0x8000 of it to the global totals   */
/* at once (an LDS subtract and a global add): each half only grows      */
/* between that add and the subtract, so neither half wraps however the  */
/* rows fall, and the block adds what is left at its end.  Bindings,     */
/* push and words as aggregate_dense (binding 2 several keys), E = 1;    */
/* global layout count u64[R] | sum u64[R] | min u32[R] | max u32[R].    */
/* ------------------------------------------------------------------ */
#define QK_NARROW_WG 1024
template <bool SUM, bool MM, bool TERMS>
__device__ __forceinline__ void qk_dense_narrow(const QK_Args& a)
{
  constexpr int WG = QK_NARROW_WG, U = QK_UNROLL;
  extern __shared__ qk_u32 lds_n[];
  const qk_u64 rows = qk_push(a, QK_P_ROWS);
  const qk_u64 r0 = qk_push(a, QK_P_ROW_BEGIN);
  const qk_u64 r1 = r0 + rows;
  const qk_u32 slots = a.words[QK_W_ARGS] >> 8;
  const qk_u32 types = a.words[QK_W_TYPES];
  const qk_u32 key_slot = a.words[QK_W_KEY] & 0xffu;
  const qk_u32 key_type = (a.words[QK_W_KEY] >> 8) & 0xfu;
  const qk_u32 key_enc = (a.words[QK_W_KEY] >> 12) & 3u;
  const qk_u32 key_min = a.words[QK_W_KEY_MIN];
  const qk_u32 range = a.words[QK_W_RANGE];
  const QkLeafSet L = qk_load_leaves<TERMS>(qk_ptr<qk_u32>(a, QK_B_LEAVES), (qk_u32)qk_push(a, QK_P_LEAVES));
  const qk_u32* kd = a.ptr[2] ? qk_ptr<qk_u32>(a, 2) : nullptr;
  const qk_u32 nkeys = kd ? kd[0] : 1u;
  qk_u32* l_cc = lds_n;
  qk_u32* l_lo = l_cc + range;
  qk_u32* l_mn = l_lo + (SUM ? range : 0u);
  qk_u32* l_mx = l_mn + (MM ? range : 0u);
  for (qk_u32 i = threadIdx.x; i < range; i += WG) {
    l_cc[i] = 0u;
    if (SUM) l_lo[i] = 0u;
    if (MM) {
      l_mn[i] = 0xffffffffu;
      l_mx[i] = 0u;
    }
  }
  __syncthreads();
  qk_u64* g_cnt = qk_wptr<qk_u64>(a, 1);
  qk_u64* g_sum = g_cnt + range;
  qk_u32* g_min = (qk_u32*)(g_sum + range);
  qk_u32* g_max = g_min + range;
  const qk_u32 aenc = (types >> 4) & 3u;
  const qk_u64 cp = a.ptr[QK_B_COL0 + (slots & 0xfu)];
  /* the block's contiguous quad range, whole steps of WG x U quads */
  qk_u64 b0, b1;
  {
    const qk_u64 q0 = r0 / 4, q1 = (r1 + 3) / 4, step = (qk_u64)WG * U;
    qk_u64 per = (q1 - q0 + gridDim.x - 1) / gridDim.x;
    per = (per + step - 1) / step * step;
    b0 = q0 + per * blockIdx.x;
    if (b0 > q1) b0 = q1;
    b1 = b0 + per < q1 ? b0 + per : q1;
  }
  for (qk_u64 s = b0; s < b1; s += (qk_u64)WG * U) {
    qk_u64 q[U];
    qk_u32 m[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const qk_u64 qq = s + (qk_u64)j * WG + threadIdx.x;
      const bool in = qq < b1;
      qk_u32 v = in ? qk_quad_valid(qq, r0, r1) : 0u;
      if (v && qk_quad_skipped(a, qq)) v = 0u;
      m[j] = v;
      q[j] = in ? qq : b1 - 1;
    }
    if (L.n) qk_leaves_step<TERMS>(a, L, q, m);
    qk_u32 g[U][4];
    if (nkeys == 1) {
#pragma unroll
      for (int j = 0; j < U; j++) {
        const qk_u32x4 kv = key_enc == 0 ? qk_ld(qk_ptr<qk_u32x4>(a, QK_B_COL0 + key_slot) + q[j])
                                         : qk_quad_codes(a.ptr[QK_B_COL0 + key_slot], key_enc, q[j]);
        const qk_u32 w[4] = {kv.x, kv.y, kv.z, kv.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          g[j][k] = qk_code32(w[k], key_type) - key_min;
          if (g[j][k] >= range) m[j] &= ~(1u << k);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < U; j++) g[j][0] = g[j][1] = g[j][2] = g[j][3] = 0u;
      for (qk_u32 kk = 0; kk < nkeys; kk++) {
        const qk_u32 d = kd[1 + 3 * kk], kmin = kd[2 + 3 * kk], kr = kd[3 + 3 * kk];
        const qk_u32 kt = (d >> 8) & 0xfu, ke = (d >> 12) & 3u;
        const qk_u64 kp = a.ptr[QK_B_COL0 + (d & 0xffu)];
#pragma unroll
        for (int j = 0; j < U; j++) {
          const qk_u32x4 x = ke == 0 ? qk_ld(qk_gptr<const qk_u32x4>(kp) + q[j]) : qk_quad_codes(kp, ke, q[j]);
          const qk_u32 w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const qk_u32 dg = qk_code32(w[k], kt) - kmin;
            if (dg >= kr) m[j] &= ~(1u << k);
            g[j][k] = g[j][k] * kr + dg;
          }
        }
      }
#pragma unroll
      for (int j = 0; j < U; j++)
#pragma unroll
        for (int k = 0; k < 4; k++)
          if (g[j][k] >= range) m[j] &= ~(1u << k);
    }
    qk_u32x4 x[U];
#pragma unroll
    for (int j = 0; j < U; j++) x[j] = aenc == 0 ? qk_ld(qk_gptr<const qk_u32x4>(cp) + q[j]) : qk_quad_codes(cp, aenc, q[j]);
#pragma unroll
    for (int j = 0; j < U; j++) {
      const qk_u32 w[4] = {x[j].x, x[j].y, x[j].z, x[j].w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (!(m[j] & (1u << k))) continue;
        const qk_u32 gi = g[j][k], v = w[k];
        qk_u32 inc = 1u;
        if (SUM) {
          const qk_u32 old = atomicAdd(&l_lo[gi], v);
          if (old + v < old) inc += 0x10000u;
        }
        const qk_u32 o = atomicAdd(&l_cc[gi], inc);
        if ((o & 0xffffu) == 0x7fffu) {
          atomicSub(&l_cc[gi], 0x8000u);
          atomicAdd(&g_cnt[gi], 0x8000ULL);
        }
        if (SUM && (inc >> 16) && (o >> 16) == 0x7fffu) {
          atomicSub(&l_cc[gi], 0x80000000u);
          atomicAdd(&g_sum[gi], 0x8000ULL << 32);
        }
        if (MM) {
          atomicMin(&l_mn[gi], v);
          atomicMax(&l_mx[gi], v);
        }
      }
    }
  }
  __syncthreads();
  for (qk_u32 i = threadIdx.x; i < range; i += WG) {
    const qk_u32 c = l_cc[i];
    if (c & 0xffffu) atomicAdd(&g_cnt[i], (qk_u64)(c & 0xffffu));
    if (SUM) {
      const qk_u64 add = ((qk_u64)(c >> 16) << 32) + l_lo[i];
      if (add) atomicAdd(&g_sum[i], add);
    }
    if (MM && l_mn[i] <= l_mx[i]) {
      atomicMin(&g_min[i], l_mn[i]);
      atomicMax(&g_max[i], l_mx[i]);
    }
  }
}

#define QK_NARROW_KERNEL(name, SUM, MM)                                                                   \
  extern "C" __global__ void __launch_bounds__(QK_NARROW_WG) k_aggregate_dense_narrow_##name(QK_Args a)  \
  {                                                                                                       \
    qk_dense_narrow<SUM, MM, false>(a);                                                                   \
  }                                                                                                       \
  extern "C" __global__ void __launch_bounds__(QK_NARROW_WG) k_aggregate_dense_narrow_##name##_or(QK_Args a) \
  {                                                                                                       \
    qk_dense_narrow<SUM, MM, true>(a);                                                                    \
  }
QK_NARROW_KERNEL(s, true, false)
QK_NARROW_KERNEL(m, false, true)
QK_NARROW_KERNEL(sm, true, true)
#undef QK_NARROW_KERNEL

/* ------------------------------------------------------------------ */
/* join_build: open-addressing table of the dim keys on the device.     */
/*  binding 0 keys (u32), 1 dim group column (u32/i32, payload mode 1), */
/*  2 table (u64 slots, 0 = empty, entry = (payload+1) << 32 | key),    */
/*  3 flags (u32[2]: [0] duplicate key seen, [1] table full),           */
/*  4 row list (u64 dim row ids from a scan, or none: all rows)         */
/*  push 0 rows to insert, words[2] log2 slots, words[3] payload mode   */
/*  (0 dim row, 1 group code = code(group) - words[4], 2 zero),          */
/*  words[5] group column type                                          */
/* Slot of a key: (key * 0x9E3779B1) >> (32 - log2 slots), linear probe  */
/* (the hashing of join_probe_native.comp / query_exec.c:8149-8191).    */
/* ------------------------------------------------------------------ */
extern "C" __global__ void LB k_join_build(QK_Args a)
{
  const qk_u64 n = qk_push(a, 0);
  const qk_u32 lg = a.words[QK_W_LOG2];
  const qk_u32 pmode = a.words[QK_W_PMODE];
  const qk_u32 gmin = a.words[QK_W_GMIN];
  const qk_u32 gtype = a.words[QK_W_GTYPE];
  const qk_u64 mask = (1ULL << lg) - 1;
  const qk_u32* keys = qk_ptr<qk_u32>(a, 0);
  const qk_u32* groups = qk_ptr<qk_u32>(a, 1);
  const qk_u64* list = a.ptr[4] ? qk_ptr<qk_u64>(a, 4) : nullptr;
  qk_u64* table = qk_wptr<qk_u64>(a, 2);
  qk_u32* flags = qk_wptr<qk_u32>(a, 3);
  for (qk_u64 i = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; i < n; i += (qk_u64)gridDim.x * QK_WG) {
    qk_u32 row = list ? (qk_u32)list[i] : (qk_u32)i;
    qk_u32 key = keys[row];
    qk_u32 pay = pmode == 0 ? row : (pmode == 1 ? qk_code32(groups[row], gtype) - gmin : 0u);
    qk_u64 entry = ((qk_u64)(pay + 1u) << 32) | key;
    qk_u64 slot = lg ? ((qk_u64)(key * 0x9E3779B1u) >> (32 - lg)) : 0;
    qk_u64 probes = 0;
    for (;;) {
      qk_u64 prev = atomicCAS(&table[slot], 0ULL, entry);
      if (prev == 0ULL) break;
      if ((qk_u32)prev == key) {
        atomicOr(&flags[0], 1u);
        break;
      }
      slot = (slot + 1) & mask;
      if (++probes > mask) {
        atomicOr(&flags[1], 1u);
        break;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* join_probe: probe the fact key of every row and aggregate.           */
/*  binding 1 accumulators (layout of aggregate_dense, R = groups),     */
/*  2 table, 5 skip, 6 fact key column, 7.. arg column slots            */
/*  push 0 rows, 5 first row; words[2] log2 slots; words[12]/[13] args  */
/*  as aggregate_dense; words[11] = R; words[14] = flags << 16          */
/*  (1 sums, 2 min/max, 4 payload is a dim row: global atomics into     */
/*  per-dim-row totals instead of LDS tables)                           */
/* ------------------------------------------------------------------ */
/* payload+1 of `key` walking from `slot`, 0 when absent */
__device__ __forceinline__ qk_u32 qk_probe_from(const qk_u64* table, qk_u64 mask, qk_u64 slot, qk_u32 key)
{
  for (qk_u64 p = 0; p < mask; p++) {
    qk_u64 e = table[slot];
    if (e == 0ULL) return 0u;
    if ((qk_u32)e == key) return (qk_u32)(e >> 32);
    slot = (slot + 1) & mask;
  }
  return 0u;
}

extern "C" __global__ void LB k_join_probe(QK_Args a)
{
  extern __shared__ qk_u64 lds_u64[];
  const qk_u64 rows = qk_push(a, QK_P_ROWS);
  const qk_u64 r0 = qk_push(a, QK_P_ROW_BEGIN);
  const qk_u64 r1 = r0 + rows;
  const qk_u32 lg = a.words[QK_W_LOG2];
  const qk_u64 mask = (1ULL << lg) - 1;
  const qk_u32 nargs = a.words[QK_W_ARGS] & 0xffu;
  const qk_u32 slots = a.words[QK_W_ARGS] >> 8;
  const qk_u32 types = a.words[QK_W_TYPES];
  const qk_u32 range = a.words[QK_W_RANGE];
  const qk_u32 flags = a.words[QK_W_KEY] >> 16;
  const bool per_row = (flags & 4u) != 0;
  const qk_u64* table = qk_ptr<qk_u64>(a, 2);
  qk_u64* g_cnt = qk_wptr<qk_u64>(a, 1);
  qk_u64* g_sum = g_cnt + range;
  qk_u32* g_min = (qk_u32*)(g_sum + (qk_u64)nargs * range);
  qk_u32* g_max = g_min + (qk_u64)nargs * range;

  qk_u64* l_sum = lds_u64;
  qk_u32* l_cnt = (qk_u32*)(lds_u64 + ((flags & 1u) ? (qk_u64)nargs * range : 0));
  qk_u32* l_min = l_cnt + range;
  qk_u32* l_max = l_min + ((flags & 2u) ? (qk_u64)nargs * range : 0);
  if (!per_row) {
    const qk_u32 sum_words = (flags & 1u) ? nargs * range : 0u;
    const qk_u32 mm_words = (flags & 2u) ? nargs * range : 0u;
    for (qk_u32 i = threadIdx.x; i < sum_words; i += QK_WG) l_sum[i] = 0;
    for (qk_u32 i = threadIdx.x; i < range; i += QK_WG) l_cnt[i] = 0;
    for (qk_u32 i = threadIdx.x; i < mm_words; i += QK_WG) {
      l_min[i] = 0xffffffffu;
      l_max[i] = 0u;
    }
    __syncthreads();
  }
  const qk_u32x4* keys = qk_ptr<qk_u32x4>(a, QK_B_COL0);
  qk_u64 b0, b1;
  qk_block_range(r0 / 4, (r1 + 3) / 4, b0, b1);
  for (qk_u64 s = b0; s < b1; s += (qk_u64)QK_WG * QK_UNROLL) {
    qk_u64 qs[QK_UNROLL];
    qk_u32 m[QK_UNROLL];
    qk_u32x4 kv[QK_UNROLL];
    qk_step_quads(a, s, b1, r0, r1, qs, m);
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) kv[j] = qk_ld(keys + qs[j]);
    /* the home slots of all 16 keys are loaded together; only keys whose
     * home slot holds another key walk on (linear probing) */
    qk_u32 pay[QK_UNROLL][4];
    qk_u64 home[QK_UNROLL][4];
    qk_u64 e[QK_UNROLL][4];
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) {
      qk_u32 w[4] = {kv[j].x, kv[j].y, kv[j].z, kv[j].w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        home[j][k] = lg ? ((qk_u64)(w[k] * 0x9E3779B1u) >> (32 - lg)) : 0;
        e[j][k] = table[home[j][k]];
      }
    }
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) {
      qk_u32 w[4] = {kv[j].x, kv[j].y, kv[j].z, kv[j].w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        qk_u32 p = 0u;
        if (m[j] & (1u << k)) {
          qk_u64 ek = e[j][k];
          if (ek != 0ULL) {
            if ((qk_u32)ek == w[k]) p = (qk_u32)(ek >> 32);
            else p = qk_probe_from(table, mask, (home[j][k] + 1) & mask, w[k]);
          }
        }
        pay[j][k] = p;
      }
    }
    qk_u32 hit[QK_UNROLL];
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) {
      hit[j] = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) hit[j] |= (pay[j][k] && pay[j][k] - 1u < range) ? (1u << k) : 0u;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (!(hit[j] & (1u << k))) continue;
        qk_u32 g = pay[j][k] - 1u;
        if (per_row) atomicAdd(&g_cnt[g], 1ULL);
        else atomicAdd(&l_cnt[g], 1u);
      }
    }
    for (qk_u32 c = 0; c < nargs; c++) {
      qk_u32 at = (types >> (8 * c)) & 0xffu;
      const qk_u32x4* col = qk_ptr<qk_u32x4>(a, QK_B_COL0 + ((slots >> (4 * c)) & 0xfu));
      qk_u32x4 x[QK_UNROLL];
#pragma unroll
      for (int j = 0; j < QK_UNROLL; j++) x[j] = qk_ld(col + qs[j]);
#pragma unroll
      for (int j = 0; j < QK_UNROLL; j++) {
        qk_u32 w[4] = {x[j].x, x[j].y, x[j].z, x[j].w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          if (!(hit[j] & (1u << k))) continue;
          qk_u32 g = pay[j][k] - 1u;
          qk_u64 add = at == QK_T_I32 ? (qk_u64)(qk_s64)(qk_s32)w[k] : (qk_u64)w[k];
          qk_u32 code = qk_code32(w[k], at);
          if (per_row) {
            if (flags & 1u) atomicAdd(&g_sum[(qk_u64)c * range + g], add);
            if (flags & 2u) {
              atomicMin(&g_min[(qk_u64)c * range + g], code);
              atomicMax(&g_max[(qk_u64)c * range + g], code);
            }
          } else {
            if (flags & 1u) atomicAdd(&l_sum[c * range + g], add);
            if (flags & 2u) {
              atomicMin(&l_min[c * range + g], code);
              atomicMax(&l_max[c * range + g], code);
            }
          }
        }
      }
    }
  }
  if (per_row) return;
  __syncthreads();
  for (qk_u32 i = threadIdx.x; i < range; i += QK_WG) {
    qk_u32 cnt = l_cnt[i];
    if (!cnt) continue;
    atomicAdd(&g_cnt[i], (qk_u64)cnt);
    for (qk_u32 c = 0; c < nargs; c++) {
      if (flags & 1u) atomicAdd(&g_sum[(qk_u64)c * range + i], l_sum[c * range + i]);
      if (flags & 2u) {
        atomicMin(&g_min[(qk_u64)c * range + i], l_min[c * range + i]);
        atomicMax(&g_max[(qk_u64)c * range + i], l_max[c * range + i]);
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* radix-partitioned join (large dimension tables).                     */
/* A key's hash h = fmix32(key) (a bijection of u32) picks its partition */
/* p = h >> (32 - pbits) and its slot inside the partition's table from */
/* the next sbits bits.  Fact rows are scattered into partition order   */
/* (hist -> offsets -> scatter), then one or more blocks per partition  */
/* load that partition's table into LDS and probe it from there, so no  */
/* probe leaves the CU.                                                 */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ qk_u32 qk_fmix32(qk_u32 h)
{
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

/* qk_fmix32(h) with its top 16 bits exact: the last step (h ^= h >> 16)
 * leaves them as they are, so a partition index (at most 12 bits, from
 * the top) needs only the steps before it */
__device__ __forceinline__ qk_u32 qk_fmix32_top16(qk_u32 h)
{
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  return h;
}

/* partition tables: buckets of QK_BUCKET slots (one 16-byte LDS read) */
#define QK_BUCKET 4

/* join_part_build: binding 0 dim keys, 1 dim group column, 2 partition
 * tables, 3 flags (u32[2]: duplicate, full), 4 row list (u64 ids,
 * optional); push 0 rows; words PMODE/GMIN/GTYPE as join_build, PBITS
 * (>= 1), SBITS.
 * With h = fmix32(key) (a bijection of u32), the partition is the top PBITS
 * bits of h and the home slot the next SBITS bits.  Inside partition p the
 * key is fully given by the low 32-PBITS bits of h, so a slot stores the key
 * word kw = (h & low) | 0x80000000 (never 0: 0 marks an empty slot) in
 * keys u32[P][S], and the payload (dim row or group code) in pays u32[P][S]
 * right after.  The slots form aligned 4-slot buckets; a key goes to the
 * first bucket from its home slot's bucket on with an empty slot, so a
 * lookup that meets a bucket with an empty slot can stop there. */
/* Packed form (PMODE bit 8, group payloads < 8192, PBITS >= 6): one u32
 * per slot, key field << 13 | group code, the key field 1 << 18 (occupied)
 * | d << RB | the low RB = 32 - PBITS - (SBITS - 2) bits of h, d the
 * bucket distance from the home bucket (< 2^(18 - RB)); h's partition and
 * home bucket bits are the slot's position less d.  A lookup subtracts the
 * expected key field << 13 from each slot word: the match alone leaves a
 * value below 8192, its group code. */
extern "C" __global__ void LB k_join_part_build(QK_Args a)
{
  const qk_u64 n = qk_push(a, 0);
  const qk_u32 pmode = a.words[QK_W_PMODE] & 0xffu, gmin = a.words[QK_W_GMIN], gtype = a.words[QK_W_GTYPE];
  const bool packed = (a.words[QK_W_PMODE] >> 8) & 1u;
  const qk_u32 pbits = a.words[QK_W_PBITS], sbits = a.words[QK_W_SBITS];
  const qk_u32 smask = (1u << sbits) - 1u;
  const qk_u32 low = 0xffffffffu >> pbits;
  const qk_u32* keys = qk_ptr<qk_u32>(a, 0);
  const qk_u32* groups = qk_ptr<qk_u32>(a, 1);
  const qk_u64* list = a.ptr[4] ? qk_ptr<qk_u64>(a, 4) : nullptr;
  qk_u32* tkeys = qk_wptr<qk_u32>(a, 2);
  qk_u32* tpays = tkeys + ((qk_u64)1 << (pbits + sbits));
  qk_u32* flags = qk_wptr<qk_u32>(a, 3);
  for (qk_u64 i = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; i < n; i += (qk_u64)gridDim.x * QK_WG) {
    qk_u32 row = list ? (qk_u32)list[i] : (qk_u32)i;
    qk_u32 key = keys[row];
    qk_u32 pay = pmode == 0 ? row : (pmode == 1 ? qk_code32(groups[row], gtype) - gmin : 0u);
    qk_u32 h = qk_fmix32(key);
    qk_u32 kw = (h & low) | 0x80000000u;
    qk_u64 tbase = (qk_u64)(h >> (32 - pbits)) << sbits;
    if (packed) {
      const qk_u32 rb = 32u - pbits - (sbits - 2u);
      const qk_u32 rem = h & ((1u << rb) - 1u);
      const qk_u32 nbk = 1u << (sbits - 2u), hb = (h >> rb) & (nbk - 1u);
      const qk_u32 maxd = 1u << (18u - rb);
      for (qk_u32 t = 0;; t++) {
        const qk_u32 d = t >> 2;
        if (d >= maxd || d >= nbk) {
          atomicOr(&flags[1], 1u);
          break;
        }
        const qk_u32 slot = (((hb + d) & (nbk - 1u)) << 2) + ((rem + t) & 3u);
        const qk_u32 kf = 0x40000u | (d << rb) | rem;
        qk_u32 prev = atomicCAS(&tkeys[tbase + slot], 0u, (kf << 13) | (pay & 0x1fffu));
        if (prev == 0u) break;
        if ((prev >> 13) == kf) {
          atomicOr(&flags[0], 1u);
          break;
        }
      }
      continue;
    }
    /* probe i takes bucket i / 4 after the home slot's bucket, and in it
     * the slot (home + i) mod 4: the home bucket's slots are all tried
     * before any later bucket's, starting at the home slot so that keys of
     * one bucket do not all contend for its first slot */
    const qk_u32 home = (h >> (32 - pbits - sbits)) & smask;
    const qk_u32 hb = home & ~(QK_BUCKET - 1u);
    for (qk_u32 i = 0;; i++) {
      if (i > smask) {
        atomicOr(&flags[1], 1u);
        break;
      }
      const qk_u32 slot = ((hb + (i & ~(QK_BUCKET - 1u))) & smask) + ((home + i) & (QK_BUCKET - 1u));
      qk_u32 prev = atomicCAS(&tkeys[tbase + slot], 0u, kw);
      if (prev == 0u) {
        tpays[tbase + slot] = pay;
        break;
      }
      if (prev == kw) {
        atomicOr(&flags[0], 1u);
        break;
      }
    }
  }
}

/* join_part_hist: binding 6 fact key column, 2 histogram out (u32
 * [blocks][P]); push 0 rows, 5 first row; words PBITS.  The block ranges
 * are those of qk_block_range, shared with join_part_scatter; the WG-thread
 * form walks them in steps of WG x U quads (1024 quads for both forms). */
template <int WG, int U>
__device__ __forceinline__ void qk_part_hist(const QK_Args& a)
{
  static_assert(WG * U == QK_WG * QK_UNROLL, "block ranges step in 1024 quads");
  extern __shared__ qk_u64 lds_u64[];
  qk_u32* hist = (qk_u32*)lds_u64;
  const qk_u64 rows = qk_push(a, QK_P_ROWS);
  const qk_u64 r0 = qk_push(a, QK_P_ROW_BEGIN);
  const qk_u64 r1 = r0 + rows;
  const qk_u32 pbits = a.words[QK_W_PBITS];
  const qk_u32 P = 1u << pbits;
  for (qk_u32 i = threadIdx.x; i < P; i += WG) hist[i] = 0;
  __syncthreads();
  const qk_u32x4* keys = qk_ptr<qk_u32x4>(a, QK_B_COL0);
  qk_u64 b0, b1;
  qk_block_range(r0 / 4, (r1 + 3) / 4, b0, b1);
  for (qk_u64 s = b0; s < b1; s += (qk_u64)WG * U) {
    qk_u32 m[U];
    qk_u32x4 kv[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const qk_u64 qq = s + (qk_u64)j * WG + threadIdx.x;
      const bool in = qq < b1;
      qk_u32 v = in ? qk_quad_valid(qq, r0, r1) : 0u;
      if (v && qk_quad_skipped(a, qq)) v = 0u;
      m[j] = v;
      kv[j] = qk_ld(keys + (in ? qq : b1 - 1));
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      qk_u32 w[4] = {kv[j].x, kv[j].y, kv[j].z, kv[j].w};
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (m[j] & (1u << k)) atomicAdd(&hist[pbits ? (qk_fmix32_top16(w[k]) >> (32 - pbits)) : 0u], 1u);
    }
  }
  __syncthreads();
  qk_u32* out = qk_wptr<qk_u32>(a, 2) + (qk_u64)blockIdx.x * P;
  for (qk_u32 i = threadIdx.x; i < P; i += WG) out[i] = hist[i];
}

extern "C" __global__ void LB k_join_part_hist(QK_Args a) { qk_part_hist<QK_WG, QK_UNROLL>(a); }
extern "C" __global__ void __launch_bounds__(512) k_join_part_hist_512(QK_Args a) { qk_part_hist<512, 2>(a); }

/* join_part_offsets: one block per partition p: exclusive scan over the
 * blocks of hist[b][p] -> off[b][p]; total[p].  binding 2 hist, 3 off,
 * 4 totals; push 0 = hist blocks; words PBITS */
extern "C" __global__ void LB k_join_part_offsets(QK_Args a)
{
  __shared__ qk_u32 s_scan[QK_WG / 64];
  const qk_u32 nb = (qk_u32)qk_push(a, 0);
  const qk_u32 P = 1u << a.words[QK_W_PBITS];
  const qk_u32 p = blockIdx.x;
  if (p >= P) return;
  const qk_u32* hist = qk_ptr<qk_u32>(a, 2);
  qk_u32* off = qk_wptr<qk_u32>(a, 3);
  qk_u32 running = 0;
  for (qk_u32 c = 0; c < nb; c += QK_WG) {
    qk_u32 b = c + threadIdx.x;
    qk_u32 v = b < nb ? hist[(qk_u64)b * P + p] : 0u;
    qk_u32 total = 0;
    qk_u32 pre = qk_block_excl_scan_u32(v, s_scan, total);
    if (b < nb) off[(qk_u64)b * P + p] = running + pre;
    running += total;
  }
  if (threadIdx.x == 0) qk_wptr<qk_u32>(a, 4)[p] = running;
}

/* exclusive scan of the P partition totals into LDS (base[p]); returns the
 * grand total.  Every block of scatter/probe recomputes it (P <= 4096). */
__device__ __forceinline__ qk_u64 qk_part_bases(const qk_u32* totals, qk_u32 P, qk_u64* base, qk_u32* s_scan)
{
  qk_u64 running = 0;
  for (qk_u32 c = 0; c < P; c += QK_WG) {
    qk_u32 i = c + threadIdx.x;
    qk_u32 v = i < P ? totals[i] : 0u;
    qk_u32 total = 0;
    qk_u32 pre = qk_block_excl_scan_u32(v, s_scan, total);
    if (i < P) base[i] = running + pre;
    running += total;
  }
  __syncthreads();
  return running;
}

/* join_part_scatter: binding 6 fact key, 7.. arg columns (slots in
 * QK_W_ARGS as join_probe), 3 off, 4 totals, 2 out rows (u32, ROWW words
 * per row: fmix32(key) then args; the buffer's last row is a discard row
 * that no partition owns); push 0 rows, 5 first row; words PBITS, ROWW.
 * Each step's 4096 rows are sorted by partition in LDS (counting sort),
 * then written out in partition runs so consecutive threads store
 * consecutive rows of one partition.  A sorted row's partition is the top
 * bits of its first word, so no partition array is kept.
 * LDS: next u64[P] (the block's next output row per partition) |
 *      dst u64[P] (output row of sorted position 0 of the partition) |
 *      cnt u32[P] | start u32[P] | buf u32[4096 * ROWW] */
#define QK_SCATTER_STEP (QK_WG * QK_UNROLL * 4)
template <int W>
struct QkRowW {
  qk_u32 w[W];
};

template <int NA>
__device__ __forceinline__ void qk_part_scatter(const QK_Args& a)
{
  extern __shared__ qk_u64 lds_u64[];
  __shared__ qk_u32 s_scan[QK_WG / 64];
  constexpr bool FIXED = NA >= 0 && NA <= 2;
  const qk_u64 rows = qk_push(a, QK_P_ROWS);
  const qk_u64 r0 = qk_push(a, QK_P_ROW_BEGIN);
  const qk_u64 r1 = r0 + rows;
  const qk_u32 pbits = a.words[QK_W_PBITS];
  const qk_u32 P = 1u << pbits;
  const qk_u32 nargs = FIXED ? (qk_u32)NA : (a.words[QK_W_ARGS] & 0xffu);
  const qk_u32 W = 1u + nargs;
  const qk_u32 slots = a.words[QK_W_ARGS] >> 8;
  qk_u64* next = lds_u64;
  qk_u64* dst = next + P;
  qk_u32* cnt = (qk_u32*)(dst + P);
  qk_u32* start = cnt + P;
  qk_u32* buf = start + P;
  qk_part_bases(qk_ptr<qk_u32>(a, 4), P, next, s_scan);
  {
    const qk_u32* off = qk_ptr<qk_u32>(a, 3) + (qk_u64)blockIdx.x * P;
    for (qk_u32 i = threadIdx.x; i < P; i += QK_WG) {
      next[i] += off[i];
      cnt[i] = 0;
    }
  }
  qk_u32* out = qk_wptr<qk_u32>(a, 2);
  /* rows [0, cap) belong to the partitions; row cap takes discarded stores */
  const qk_u64 nrows = a.bytes[2] / (4ULL * W);
  const qk_u64 cap = nrows ? nrows - 1 : 0;
  const qk_u32 pshift = 32 - pbits;
  const qk_u32x4* keys = qk_ptr<qk_u32x4>(a, QK_B_COL0);
  const qk_u32x4* col0 = nargs ? qk_ptr<qk_u32x4>(a, QK_B_COL0 + (slots & 0xfu)) : keys;
  qk_u64 b0, b1;
  qk_block_range(r0 / 4, (r1 + 3) / 4, b0, b1);
  const qk_u64 qstep = (qk_u64)QK_WG * QK_UNROLL;
  /* the keys of step s+1 load while step s sorts and writes, and the first
   * arg column of step s loads at its start (it is first needed after the
   * ranking and the scan).  The loads are unconditional (past the range
   * they reread the last quad, masked off), and with the write-out's store
   * count fixed the wait for them does not also wait for this step's
   * stores.  The loop's barriers order LDS only. */
  qk_u64 nq[QK_UNROLL];
  qk_u32 nm[QK_UNROLL];
  qk_u32x4 nkv[QK_UNROLL];
  auto load_step = [&](qk_u64 s) {
    qk_step_quads(a, s, b1, r0, r1, nq, nm);
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) nkv[j] = keys[nq[j]];
  };
  if (b0 < b1) load_step(b0);
  qk_lds_barrier();
  auto step = [&](qk_u64 s) {
    qk_u64 q[QK_UNROLL];
    qk_u32 m[QK_UNROLL];
    qk_u32x4 kv[QK_UNROLL], x0[QK_UNROLL];
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) {
      q[j] = nq[j];
      m[j] = nm[j];
      kv[j] = nkv[j];
      if (nargs) x0[j] = qk_ld(col0 + q[j]);
    }
    load_step(s + qstep < b1 ? s + qstep : s);
    /* a row is stored with its key's hash h = fmix32(key) in place of the
     * key (a bijection, so nothing is lost): the probe reads the partition,
     * slot and key word straight from it */
    qk_u32 hv[QK_UNROLL][4], rank[QK_UNROLL][4];
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) {
      qk_u32 w[4] = {kv[j].x, kv[j].y, kv[j].z, kv[j].w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        hv[j][k] = qk_fmix32(w[k]);
        rank[j][k] = (m[j] & (1u << k)) ? atomicAdd(&cnt[pbits ? hv[j][k] >> pshift : 0u], 1u) : 0u;
      }
    }
    qk_lds_barrier();
    /* start[p] = exclusive scan of cnt; dst[p] = output row of sorted
     * position 0 of partition p (u64 arithmetic modulo 2^64: dst[p] + r
     * is the row of sorted position r >= start[p]); cnt reset for the next
     * step (its atomics follow a barrier that this scan's readers pass) */
    qk_u32 running = 0;
    for (qk_u32 c = 0; c < P; c += QK_WG) {
      qk_u32 i = c + threadIdx.x;
      qk_u32 v = i < P ? cnt[i] : 0u;
      qk_u32 total = 0;
      qk_u32 pre = qk_block_excl_scan_u32<true>(v, s_scan, total);
      if (i < P) {
        const qk_u32 st = running + pre;
        start[i] = st;
        dst[i] = next[i] - (qk_u64)st;
        next[i] += v;
        cnt[i] = 0;
      }
      running += total;
    }
    const qk_u32 step_rows = running;
    qk_lds_barrier();
    /* place rows into partition order: every start[] read of the step is
     * issued before the first write */
    qk_u32 idx[QK_UNROLL][4];
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++)
#pragma unroll
      for (int k = 0; k < 4; k++) idx[j][k] = start[pbits ? hv[j][k] >> pshift : 0u] + rank[j][k];
    if (FIXED && W <= 2) {
#pragma unroll
      for (int j = 0; j < QK_UNROLL; j++) {
        qk_u32 x[4] = {x0[j].x, x0[j].y, x0[j].z, x0[j].w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          if (!(m[j] & (1u << k))) continue;
          if (W == 2) ((uint2*)buf)[idx[j][k]] = make_uint2(hv[j][k], x[k]);
          else buf[idx[j][k]] = hv[j][k];
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < QK_UNROLL; j++) {
        qk_u32 x[4] = {x0[j].x, x0[j].y, x0[j].z, x0[j].w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          if (!(m[j] & (1u << k))) continue;
          buf[idx[j][k] * W] = hv[j][k];
          buf[idx[j][k] * W + 1] = x[k];
        }
      }
      for (qk_u32 c = 1; c < nargs; c++) {
        const qk_u32x4* col = qk_ptr<qk_u32x4>(a, QK_B_COL0 + ((slots >> (4 * c)) & 0xfu));
        qk_u32x4 x[QK_UNROLL];
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) x[j] = qk_ld(col + q[j]);
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) {
          qk_u32 w[4] = {x[j].x, x[j].y, x[j].z, x[j].w};
#pragma unroll
          for (int k = 0; k < 4; k++)
            if (m[j] & (1u << k)) buf[idx[j][k] * W + 1 + c] = w[k];
        }
      }
    }
    qk_lds_barrier();
    /* write out: consecutive threads take consecutive sorted rows, so a
     * partition's run goes out as contiguous stores.  Fixed shapes issue
     * every read of the step, then every store (a fixed count of stores,
     * positions past the step's rows going to the discard row). */
    if (FIXED) {
      constexpr int WF = FIXED ? NA + 1 : 1;
      constexpr int PER = QK_SCATTER_STEP / QK_WG;
      const QkRowW<WF>* rb = (const QkRowW<WF>*)buf;
      QkRowW<WF>* ro = (QkRowW<WF>*)out;
      constexpr int HALF = PER / 2;
#pragma unroll
      for (int h = 0; h < PER; h += HALF) {
        QkRowW<WF> row[HALF];
        qk_u64 d[HALF];
#pragma unroll
        for (int i = 0; i < HALF; i++) row[i] = rb[(h + i) * QK_WG + threadIdx.x];
#pragma unroll
        for (int i = 0; i < HALF; i++) {
          qk_u32 r = (qk_u32)(h + i) * QK_WG + threadIdx.x;
          /* opaque to the optimiser: otherwise the 16 loop-invariant
           * (out + r) address bases are hoisted out of the step loop and
           * held in (spilled) registers */
          asm volatile("" : "+v"(r));
          d[i] = dst[pbits ? row[i].w[0] >> pshift : 0u] + r;
          if (r >= step_rows || d[i] >= cap) d[i] = cap;
        }
#pragma unroll
        for (int i = 0; i < HALF; i++) ro[d[i]] = row[i];
      }
    } else {
      for (qk_u32 i = threadIdx.x; i < step_rows * W; i += QK_WG) {
        qk_u32 r = i / W, wd = i - r * W;
        qk_u64 dr = dst[pbits ? buf[r * W] >> pshift : 0u] + r;
        if (dr < cap) out[dr * W + wd] = buf[i];
      }
    }
  };
  /* the first step is peeled so the loop is entered, like its back edge,
   * with a step's stores in flight: the waits for the next keys then count
   * past those stores instead of draining them */
  if (b0 < b1) step(b0);
  for (qk_u64 s = b0 + qstep; s < b1; s += qstep) step(s);
}

/* instances over the arg count (rows of 1 + NA words); NA < 0 reads it */
/* four blocks per CU fit the LDS of the 1- and 2-word rows at P <= 256;
 * the register budget is held to match (128 VGPRs: 4 waves per SIMD) */
#define LB_SCATTER LB __attribute__((amdgpu_waves_per_eu(4)))
extern "C" __global__ void LB k_join_part_scatter(QK_Args a) { qk_part_scatter<-1>(a); }
extern "C" __global__ void LB_SCATTER k_join_part_scatter_0(QK_Args a) { qk_part_scatter<0>(a); }
extern "C" __global__ void LB_SCATTER k_join_part_scatter_1(QK_Args a) { qk_part_scatter<1>(a); }
extern "C" __global__ void LB k_join_part_scatter_2(QK_Args a) { qk_part_scatter<2>(a); }

/* join_part_scatter_wc_<NA>: join_part_scatter for rows of 1 or 2 words
 * (NA = 0, 1) that writes the partitions' rows as whole aligned 128-byte
 * lines.  A line left partly filled at the end of a step is held in LDS
 * (the partition's carry) and written when a later step completes it, or
 * at the block's end; written out partly filled between steps, such lines
 * cost HBM write bandwidth.  The block's first line of each partition may
 * hold rows of the previous block: those rows are never written here.
 * Bindings, push and words as join_part_scatter; 512-thread blocks over
 * the block ranges of join_part_hist (4096-row steps, 8 rows a thread);
 * row indices < 2^32.
 * LDS: next u64[P] | cnt, start, nf, vc u32[P] | lines QkWcLine[P +
 *      4096 / CH] | carry u32[P * 32] | buf u32[4096 * ROWW], where CH =
 *      32 / ROWW rows make one line. */
#define QK_WC_WG 512
struct QkWcLine {
  qk_u32 dst;   /* output line index (row / CH) */
  qk_u32 p;     /* partition */
  qk_s32 bo;    /* buf row of the line's row 0 (rows below c are carried) */
  qk_u32 cs;    /* c (carried rows leading the line) | skip << 16 */
};

template <int NA>
__device__ __forceinline__ void qk_part_scatter_wc(const QK_Args& a)
{
  constexpr int WG = QK_WC_WG, U = 2;
  constexpr qk_u32 W = 1 + NA;
  constexpr qk_u32 CH = 32 / W, CHB = W == 1 ? 5 : 4;
  constexpr qk_u32 RT = 4 / W; /* rows in 16 bytes */
  static_assert(NA == 0 || NA == 1, "rows of 1 or 2 words");
  using Row = QkRowW<W>;
  extern __shared__ qk_u64 lds_u64[];
  __shared__ qk_u32 s_scan[WG / 64];
  const qk_u64 rows = qk_push(a, QK_P_ROWS);
  const qk_u64 r0 = qk_push(a, QK_P_ROW_BEGIN);
  const qk_u64 r1 = r0 + rows;
  const qk_u32 pbits = a.words[QK_W_PBITS];
  const qk_u32 P = 1u << pbits;
  const qk_u32 pshift = 32 - pbits;
  const qk_u32 slots = a.words[QK_W_ARGS] >> 8;
  qk_u64* next = lds_u64;
  qk_u32* cnt = (qk_u32*)(next + P);
  qk_u32* start = cnt + P;
  qk_u32* nf = start + P;
  qk_u32* vc = nf + P;
  QkWcLine* lines = (QkWcLine*)(vc + P);
  Row* carry = (Row*)(lines + P + 4096 / CH);
  Row* buf = carry + (qk_u64)P * CH;
  Row* out = (Row*)qk_wptr<qk_u32>(a, 2);
  const qk_u64 cap = a.bytes[2] / (4ULL * W);
  {
    /* next[p] = the block's first row of partition p; its leading rows in
     * that row's line belong to earlier blocks (vc[p] of them) */
    const qk_u32* totals = qk_ptr<qk_u32>(a, 4);
    const qk_u32* off = qk_ptr<qk_u32>(a, 3) + (qk_u64)blockIdx.x * P;
    qk_u64 running = 0;
    for (qk_u32 c0 = 0; c0 < P; c0 += WG) {
      const qk_u32 i = c0 + threadIdx.x;
      const qk_u32 v = i < P ? totals[i] : 0u;
      qk_u32 total = 0;
      const qk_u32 pre = qk_block_excl_scan_wg<WG, true>(v, s_scan, total);
      if (i < P) {
        const qk_u64 n0 = running + pre + off[i];
        next[i] = n0;
        vc[i] = (qk_u32)(n0 & (CH - 1));
        cnt[i] = 0;
      }
      running += total;
    }
  }
  const qk_u32x4* keys = qk_ptr<qk_u32x4>(a, QK_B_COL0);
  const qk_u32x4* col0 = NA ? qk_ptr<qk_u32x4>(a, QK_B_COL0 + (slots & 0xfu)) : keys;
  qk_u64 b0, b1;
  qk_block_range(r0 / 4, (r1 + 3) / 4, b0, b1);
  const qk_u64 qstep = (qk_u64)WG * U;
  qk_u32 nm[U];
  qk_u32x4 nkv[U], nx[U];
  /* quads s + j * WG + tid; past the range the last quad, masked off */
  auto load_step = [&](qk_u64 s) {
#pragma unroll
    for (int j = 0; j < U; j++) {
      const qk_u64 qq = s + (qk_u64)j * WG + threadIdx.x;
      const bool in = qq < b1;
      qk_u32 v = in ? qk_quad_valid(qq, r0, r1) : 0u;
      if (v && qk_quad_skipped(a, qq)) v = 0u;
      nm[j] = v;
      const qk_u64 q = in ? qq : b1 - 1;
      nkv[j] = qk_ld(keys + q);
      if (NA) nx[j] = qk_ld(col0 + q);
    }
  };
  if (b0 < b1) load_step(b0);
  qk_lds_barrier();
  for (qk_u64 s = b0; s < b1; s += qstep) {
    qk_u32 m[U];
    qk_u32x4 kv[U], xv[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      m[j] = nm[j];
      kv[j] = nkv[j];
      xv[j] = nx[j];
    }
    load_step(s + qstep < b1 ? s + qstep : s);
    /* rank each row within its partition's rows of the step */
    qk_u32 hv[U][4], rank[U][4];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const qk_u32 w[4] = {kv[j].x, kv[j].y, kv[j].z, kv[j].w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        hv[j][k] = qk_fmix32(w[k]);
        rank[j][k] = (m[j] & (1u << k)) ? atomicAdd(&cnt[pbits ? hv[j][k] >> pshift : 0u], 1u) : 0u;
      }
    }
    qk_lds_barrier();
    /* per partition: c carried rows + cnt new rows; the whole lines among
     * them are written this step.  One scan of cnt << 16 | lines gives the
     * new rows' buf offsets and the lines' numbering (a step has at most
     * 4096 rows and P + 4096 / CH lines, so neither half overflows). */
    qk_u32 run_rows = 0, run_lines = 0;
    for (qk_u32 c0 = 0; c0 < P; c0 += WG) {
      const qk_u32 i = c0 + threadIdx.x;
      qk_u32 cn = 0, c = 0, f = 0;
      qk_u64 nx0 = 0;
      if (i < P) {
        cn = cnt[i];
        nx0 = next[i];
        c = (qk_u32)(nx0 & (CH - 1));
        f = (c + cn) >> CHB;
      }
      qk_u32 total = 0;
      const qk_u32 pre = qk_block_excl_scan_wg<WG, true>((cn << 16) | f, s_scan, total);
      if (i < P) {
        const qk_u32 st = run_rows + (pre >> 16);
        const qk_u32 l0 = run_lines + (pre & 0xffffu);
        start[i] = st;
        nf[i] = (c + cn) | (c << 16);
        const qk_u32 skip = vc[i];
        const qk_u32 d0 = (qk_u32)((nx0 - c) >> CHB);
        for (qk_u32 k = 0; k < f; k++) {
          QkWcLine L;
          L.dst = d0 + k;
          L.p = i;
          L.bo = (qk_s32)(st + k * CH) - (qk_s32)c;
          L.cs = k ? 0u : (c | (skip << 16));
          lines[l0 + k] = L;
        }
        if (f) vc[i] = 0;
        next[i] = nx0 + cn;
        cnt[i] = 0;
      }
      run_rows += total >> 16;
      run_lines += total & 0xffffu;
    }
    const qk_u32 nlines = run_lines;
    qk_lds_barrier();
    /* place the new rows in partition order */
#pragma unroll
    for (int j = 0; j < U; j++) {
      const qk_u32 x[4] = {xv[j].x, xv[j].y, xv[j].z, xv[j].w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (!(m[j] & (1u << k))) continue;
        Row r;
        r.w[0] = hv[j][k];
        if (W == 2) r.w[W - 1] = x[k];
        buf[start[pbits ? hv[j][k] >> pshift : 0u] + rank[j][k]] = r;
      }
    }
    qk_lds_barrier();
    /* write the whole lines: CH / RT consecutive threads take one line, a
     * thread RT rows (16 bytes) in one store; the store unit spends its
     * time per instruction.  A thread whose rows straddle the block's
     * leading skip or the buffer's end stores them one by one. */
    for (qk_u32 idx = threadIdx.x; idx < nlines * (CH / RT); idx += WG) {
      const QkWcLine L = lines[idx / (CH / RT)];
      const qk_u32 jj = RT * (idx % (CH / RT));
      const qk_u32 c = L.cs & 0xffffu, skip = L.cs >> 16;
      Row r[RT];
#pragma unroll
      for (qk_u32 e = 0; e < RT; e++)
        r[e] = jj + e < c ? carry[(qk_u64)L.p * CH + jj + e] : buf[L.bo + (qk_s32)(jj + e)];
      const qk_u64 d = ((qk_u64)L.dst << CHB) + jj;
      if (jj >= skip && d + RT <= cap) {
        qk_u32x4 v;
        if (W == 2) {
          v.x = r[0].w[0];
          v.y = r[0].w[W - 1];
          v.z = r[RT - 1].w[0];
          v.w = r[RT - 1].w[W - 1];
        } else {
          v.x = r[0].w[0];
          v.y = r[1 % RT].w[0];
          v.z = r[2 % RT].w[0];
          v.w = r[3 % RT].w[0];
        }
        *(qk_u32x4*)(out + d) = v;
      } else {
#pragma unroll
        for (qk_u32 e = 0; e < RT; e++)
          if (jj + e >= skip && d + e < cap) out[d + e] = r[e];
      }
    }
    qk_lds_barrier();
    /* carry the rows past each partition's last whole line, stored from
     * the registers that still hold them: a row at position c + rank of
     * its partition's c carried and n rows is past the whole lines when
     * that position is at or past n rounded down to a line (the lines
     * above have read the carried rows of the step before) */
#pragma unroll
    for (int j = 0; j < U; j++) {
      const qk_u32 x[4] = {xv[j].x, xv[j].y, xv[j].z, xv[j].w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (!(m[j] & (1u << k))) continue;
        const qk_u32 p = pbits ? hv[j][k] >> pshift : 0u;
        const qk_u32 info = nf[p];
        const qk_u32 pos = (info >> 16) + rank[j][k], fl = (info & 0xffffu) & ~(CH - 1u);
        if (pos >= fl) {
          Row r;
          r.w[0] = hv[j][k];
          if (W == 2) r.w[W - 1] = x[k];
          carry[(qk_u64)p * CH + (pos - fl)] = r;
        }
      }
    }
  }
  qk_lds_barrier();
  /* the partly filled last lines */
  for (qk_u32 t = threadIdx.x; t < P * CH; t += WG) {
    const qk_u32 p = t >> CHB, jj = t & (CH - 1);
    const qk_u64 nx0 = next[p];
    const qk_u32 c = (qk_u32)(nx0 & (CH - 1));
    const qk_u64 d = nx0 - c + jj;
    if (jj < c && jj >= vc[p] && d < cap) out[d] = carry[t];
  }
}

extern "C" __global__ void __launch_bounds__(QK_WC_WG) k_join_part_scatter_wc_0(QK_Args a) { qk_part_scatter_wc<0>(a); }
extern "C" __global__ void __launch_bounds__(QK_WC_WG) k_join_part_scatter_wc_1(QK_Args a) { qk_part_scatter_wc<1>(a); }

/* join_part_probe: blocks (partition p, split i) for i < SPLIT; loads
 * table p (join_part_build's layout) into LDS and probes its share of the
 * partitioned rows; aggregates like join_probe (flags 1 sums, 2 min/max;
 * payload = group code -> LDS group tables; flag 4 payload = dim row ->
 * global atomics).  binding 1 accumulators, 2 partition tables, 3 rows,
 * 4 totals; words PBITS (>= 1), SBITS (>= 3), ROWW, SPLIT, ARGS (arg
 * count | log2 of the group-table copies << 16), TYPES, RANGE, KEY
 * (flags << 16).
 * A lookup reads the key's home bucket and the bucket after it (two
 * 16-byte LDS reads issued together) and compares their key words: a key
 * sits between its home bucket's first slot and the first empty slot at or
 * after it, so the two buckets settle the lookup unless both are full
 * without the key, and only then do further buckets follow.
 * In group mode the LDS group tables are replicated (2^k copies, lane i
 * adding into copy i mod 2^k, each copy's stride odd so one group's copies
 * fall in different banks): lanes of a wave adding into one group then
 * mostly hit different addresses.  The copies are summed at the end.
 * The body is a template over the arg count and the aggregate flags so the
 * common shapes run without per-row branches on them; NA < 0 is the
 * generic form that reads them from the push words.  Group codes
 * (< 65536) sit in LDS as u16, dim rows as u32. */
struct QkBucket {
  qk_u32 w[QK_BUCKET];
};

__device__ __forceinline__ QkBucket qk_bucket_read(const qk_u32* tk, qk_u32 b)
{
  const qk_u32x4 k0 = *(const qk_u32x4*)(tk + b);
  return QkBucket{{k0.x, k0.y, k0.z, k0.w}};
}

__device__ __forceinline__ qk_u32 qk_bucket_match(const QkBucket& k, qk_u32 v)
{
  qk_u32 m = 0u;
#pragma unroll
  for (int i = 0; i < QK_BUCKET; i++) m |= (k.w[i] == v) ? (1u << i) : 0u;
  return m;
}

/* PK: the table in join_part_build's packed form (group mode): a bucket
 * read gives keys and group codes at once, and a row's home bucket and the
 * next are matched by subtraction and unsigned minimum (the matching slot
 * alone leaves its code, below 8192); a block whose rows' count and sum of
 * one U32 argument fit 24 + 40 bits adds both with one 64-bit LDS atomic. */
template <int NA, bool SUMS, bool MM, bool ROW, int WG, int ROWS, int G, bool PK = false>
__device__ __forceinline__ void qk_part_probe(const QK_Args& a)
{
  /* 16-byte aligned: the bucket reads are 16-byte LDS loads */
  extern __shared__ __attribute__((aligned(16))) qk_u64 lds_u64[];
  __shared__ qk_u64 s_red[WG / 64];
  constexpr bool FIXED = NA >= 0;
  const qk_u32 pbits = a.words[QK_W_PBITS], sbits = a.words[QK_W_SBITS];
  const qk_u32 P = 1u << pbits, S = 1u << sbits, smask = S - 1u;
  const qk_u32 low = 0xffffffffu >> pbits;
  const qk_u32 split = a.words[QK_W_SPLIT];
  const qk_u32 nargs = FIXED ? (qk_u32)NA : (a.words[QK_W_ARGS] & 0xffu);
  const qk_u32 W = 1u + nargs;
  const qk_u32 types = a.words[QK_W_TYPES];
  const qk_u32 range = a.words[QK_W_RANGE];
  const qk_u32 flags = a.words[QK_W_KEY] >> 16;
  const bool do_sum = FIXED ? SUMS : (flags & 1u) != 0;
  const bool do_mm = FIXED ? MM : (flags & 2u) != 0;
  const bool per_row = FIXED ? ROW : (flags & 4u) != 0;
  const qk_u32 part = blockIdx.x / split, piece = blockIdx.x % split;
  if (part >= P) return;
  /* LDS: keys u32[S] | pays (u16 group codes or u32 dim rows)[S] | then,
   * in group mode, C copies each of sums u64[E][R] | cnt u32[R] |
   * min u32[E][R] | max u32[E][R], every copy's stride odd */
  static_assert(!PK || !ROW, "the packed table holds group codes");
  qk_u32* tk = (qk_u32*)lds_u64;
  unsigned short* tp16 = (unsigned short*)(tk + S);
  qk_u32* tp32 = tk + S;
  /* packed: low bits of h below the bucket index, and the key-field step
   * of one bucket of distance */
  const qk_u32 rb = 32u - pbits - (sbits - 2u);
  const qk_u32 rmask = (1u << rb) - 1u, nbk = S >> 2;
  const qk_u32 dstep = 1u << (rb + 13u), maxd = 1u << (18u - rb);
  const qk_u32 C = per_row ? 1u : (1u << ((a.words[QK_W_ARGS] >> 16) & 0xfu));
  const qk_u32 sw = (!per_row && do_sum) ? nargs * range : 0u;
  const qk_u32 mw = (!per_row && do_mm) ? nargs * range : 0u;
  const qk_u32 sstr = sw ? (sw | 1u) : 0u, cstr = per_row ? 0u : (range | 1u), mstr = mw ? (mw | 1u) : 0u;
  qk_u64* l_sum = (qk_u64*)(PK ? (void*)(tk + S) : (per_row ? (void*)(tp32 + S) : (void*)(tp16 + S)));
  qk_u32* l_cnt = (qk_u32*)(l_sum + (qk_u64)C * sstr);
  qk_u32* l_min = l_cnt + C * cstr;
  qk_u32* l_max = l_min + C * mstr;
  const qk_u32 copy = __lane_id() & (C - 1u);
  qk_u64* my_sum = l_sum + copy * sstr;
  qk_u32* my_cnt = l_cnt + copy * cstr;
  qk_u32* my_min = l_min + copy * mstr;
  qk_u32* my_max = l_max + copy * mstr;
  {
    const qk_u32x4* gk = qk_ptr<qk_u32x4>(a, 2) + (((qk_u64)part << sbits) >> 2);
    const qk_u32x4* gp = qk_ptr<qk_u32x4>(a, 2) + ((((qk_u64)P + part) << sbits) >> 2);
    for (qk_u32 i = threadIdx.x; i < S / 4; i += WG) {
      ((qk_u32x4*)tk)[i] = gk[i];
      if (PK) continue;
      qk_u32x4 v = gp[i];
      if (per_row) {
        ((qk_u32x4*)tp32)[i] = v;
      } else {
        tp16[4 * i] = (unsigned short)v.x;
        tp16[4 * i + 1] = (unsigned short)v.y;
        tp16[4 * i + 2] = (unsigned short)v.z;
        tp16[4 * i + 3] = (unsigned short)v.w;
      }
    }
  }
  if (!per_row) {
    for (qk_u32 i = threadIdx.x; i < C * sstr; i += WG) l_sum[i] = 0;
    for (qk_u32 i = threadIdx.x; i < C * cstr; i += WG) l_cnt[i] = 0;
    for (qk_u32 i = threadIdx.x; i < C * mstr; i += WG) {
      l_min[i] = 0xffffffffu;
      l_max[i] = 0u;
    }
  }
  /* this partition's first row: the rows of partitions 0..part-1 precede
   * it (the reduction's barriers also publish the LDS fills above) */
  const qk_u32* totals = qk_ptr<qk_u32>(a, 4);
  qk_u64 before = 0;
  for (qk_u32 i = threadIdx.x; i < part; i += WG) before += totals[i];
  const qk_u64 pstart = qk_block_sum_u64_wg<WG>(before, s_red);
  const qk_u64 pend = pstart + totals[part];
  const qk_u64 plen = pend - pstart;
  const qk_u64 chunk = (plen + split - 1) / split;
  const qk_u64 rs = pstart + chunk * piece < pend ? pstart + chunk * piece : pend;
  const qk_u64 re = rs + chunk < pend ? rs + chunk : pend;
  /* one atomic for count and sum: a U32 argument whose block total stays
   * below 2^40 (its zone-map maximum, word GMIN, times the block's rows)
   * and a count below 2^24 */
  const bool pk_cs = PK && NA == 1 && SUMS && !MM && (types & 0xffu) == QK_T_U32 && re - rs < (1ull << 23) &&
                     (re - rs) * (qk_u64)a.words[QK_W_GMIN] < (1ull << 40);
  const qk_u32* rowsbuf = qk_ptr<qk_u32>(a, 3);
  qk_u64* g_cnt = qk_wptr<qk_u64>(a, 1);
  qk_u64* g_sum = g_cnt + range;
  qk_u32* g_min = (qk_u32*)(g_sum + (qk_u64)nargs * range);
  qk_u32* g_max = g_min + (qk_u64)nargs * range;
  const qk_u64 pstep = (qk_u64)WG * ROWS;
  if constexpr (PK && NA <= 1) {
    /* Rows come in 16-byte loads, RL consecutive rows a lane per load:
     * the load unit spends its time per load instruction, and with 8-byte
     * loads it was busy for nearly all of the probe.  Lane t's rows in a
     * step from r are RL * (j * WG + t) + e (e < RL); the block's rows
     * start at rs rounded down to a multiple of RL (a0), so every load is
     * 16-byte aligned, and rows before rs or at or past re are masked.  A
     * group of rows that starts before re stays inside the buffer (the host
     * leaves RL - 1 rows past the partitions' rows); a group at or past re
     * reads the group at a0 instead.  Two register sets of rows alternate:
     * one set's loads are in flight while the other is probed.  The loads
     * are unconditional: loaded under a condition, a set's values would be
     * copied into the registers that carry it round the loop, and the
     * copies would wait for the loads at once instead of when the set is
     * probed. */
    constexpr int NX = NA > 0 ? NA : 1;
    constexpr qk_u32 RL = NA > 0 ? 2u : 4u;
    static_assert(ROWS % RL == 0, "whole 16-byte groups of rows a step");
    const qk_u64 a0 = rs & ~(qk_u64)(RL - 1u);
    qk_u32 k0[ROWS], k1[ROWS], x0[ROWS][NX], x1[ROWS][NX];
    auto load_set = [&](qk_u32(&kk)[ROWS], qk_u32(&xx)[ROWS][NX], qk_u64 r) {
      const qk_u64 left = r < re ? re - r : 0u;
#pragma unroll
      for (int j = 0; j < ROWS / (int)RL; j++) {
        const qk_u32 o = RL * ((qk_u32)j * WG + threadIdx.x);
        const qk_u32x4 v = qk_ld((const qk_u32x4*)(rowsbuf + (o < left ? r + o : a0) * W));
        if constexpr (NA > 0) {
          kk[RL * j] = v.x;
          xx[RL * j][0] = v.y;
          kk[RL * j + 1] = v.z;
          xx[RL * j + 1][0] = v.w;
        } else {
          kk[RL * j] = v.x;
          kk[RL * j + 1] = v.y;
          kk[RL * j + 2] = v.z;
          kk[RL * j + 3] = v.w;
        }
      }
    };
    auto probe_set = [&](const qk_u32(&kk)[ROWS], const qk_u32(&xx)[ROWS][NX], qk_u64 r) {
      const qk_u64 left = re - r;
      const qk_u32 lim = left < pstep ? (qk_u32)left : 0xffffffffu;
      const qk_u32 lo = rs > r ? (qk_u32)(rs - r) : 0u; /* below RL, in the first step only */
      const qk_u32 tl = RL * threadIdx.x;
#pragma unroll
      for (int sb = 0; sb < ROWS; sb += G) {
        qk_u32 e0[G], home[G], d[G];
        QkBucket bk0[G], bk1[G];
#pragma unroll
        for (int j = 0; j < G; j++) {
          const qk_u32 h = kk[sb + j];
          e0[j] = (0x40000u | (h & rmask)) << 13;
          home[j] = ((h >> rb) & (nbk - 1u)) << 2;
        }
#pragma unroll
        for (int j = 0; j < G; j++) {
          bk0[j] = qk_bucket_read(tk, home[j]);
          bk1[j] = qk_bucket_read(tk, (home[j] + QK_BUCKET) & smask);
        }
#pragma unroll
        for (int j = 0; j < G; j++) {
          const qk_u32 a0 = e0[j], a1 = a0 + dstep;
          qk_u32 m = min(min(bk0[j].w[0] - a0, bk0[j].w[1] - a0), min(bk0[j].w[2] - a0, bk0[j].w[3] - a0));
          m = min(m, min(min(bk1[j].w[0] - a1, bk1[j].w[1] - a1), min(bk1[j].w[2] - a1, bk1[j].w[3] - a1)));
          if (m >= 8192u) {
            /* not in the first two buckets: on while none of the buckets
             * read had an empty slot (rare) */
            qk_u32 z = min(min(min(bk0[j].w[0], bk0[j].w[1]), min(bk0[j].w[2], bk0[j].w[3])),
                           min(min(bk1[j].w[0], bk1[j].w[1]), min(bk1[j].w[2], bk1[j].w[3])));
            qk_u32 b = (home[j] + QK_BUCKET) & smask, e = a1;
            for (qk_u32 n = 2; z != 0u && m >= 8192u && n < maxd && n < nbk; n++) {
              b = (b + QK_BUCKET) & smask;
              e += dstep;
              const QkBucket nb = qk_bucket_read(tk, b);
              m = min(m, min(min(nb.w[0] - e, nb.w[1] - e), min(nb.w[2] - e, nb.w[3] - e)));
              z = min(min(nb.w[0], nb.w[1]), min(nb.w[2], nb.w[3]));
            }
          }
          /* a miss leaves m >= 8192 >= range; rows outside [rs, re) are
           * dropped */
          const qk_u32 idx = RL * (qk_u32)((sb + j) / (int)RL) * WG + tl + (qk_u32)((sb + j) % (int)RL);
          d[j] = (m < range && idx >= lo && idx < lim) ? m : 0xffffffffu;
        }
        if (pk_cs) {
#pragma unroll
          for (int j = 0; j < G; j++)
            if (d[j] != 0xffffffffu) atomicAdd(&my_sum[d[j]], (1ULL << 40) | (qk_u64)xx[sb + j][0]);
          continue;
        }
#pragma unroll
        for (int j = 0; j < G; j++)
          if (d[j] != 0xffffffffu) atomicAdd(&my_cnt[d[j]], 1u);
#pragma unroll
        for (int c = 0; c < NA; c++) {
          const qk_u32 at = (types >> (8 * c)) & 0xffu;
#pragma unroll
          for (int j = 0; j < G; j++) {
            if (d[j] == 0xffffffffu) continue;
            const qk_u32 v = xx[sb + j][c], g = d[j];
            if (SUMS) atomicAdd(&my_sum[c * range + g], at == QK_T_I32 ? (qk_u64)(qk_s64)(qk_s32)v : (qk_u64)v);
            if (MM) {
              const qk_u32 code = qk_code32(v, at);
              atomicMin(&my_min[c * range + g], code);
              atomicMax(&my_max[c * range + g], code);
            }
          }
        }
      }
    };
    if (rs < re) load_set(k0, x0, a0);
    for (qk_u64 r = a0; r < re; r += 2 * pstep) {
      load_set(k1, x1, r + pstep);
      probe_set(k0, x0, r);
      if (r + pstep >= re) break;
      load_set(k0, x0, r + 2 * pstep);
      probe_set(k1, x1, r + pstep);
    }
  } else {
  /* ROWS rows per thread per step.  A row's key and first arg
   * load together, one whole step ahead of their use, so a step's loads are
   * in flight while the step before probes the LDS table; the rows are then
   * probed G at a time to bound the registers held per row.  Rows past
   * re clamp to rs (in bounds: the loop runs only when rs < re). */
  qk_u32 nkey[ROWS], narg[ROWS];
  auto load_rows = [&](qk_u64 r) {
#pragma unroll
    for (int j = 0; j < ROWS; j++) {
      qk_u64 x = r + (qk_u64)j * WG + threadIdx.x;
      qk_u64 y = x < re ? x : rs;
      if (W == 2) {
        const qk_u64 v = qk_ld((const qk_u64*)rowsbuf + y);
        nkey[j] = (qk_u32)v;
        narg[j] = (qk_u32)(v >> 32);
      } else {
        nkey[j] = rowsbuf[y * W];
        narg[j] = W > 1 ? rowsbuf[y * W + 1] : 0u;
      }
    }
  };
  if (rs < re) load_rows(rs);
  for (qk_u64 r = rs; r < re; r += pstep) {
    qk_u32 ckey[ROWS], carg[ROWS];
#pragma unroll
    for (int j = 0; j < ROWS; j++) {
      ckey[j] = nkey[j];
      carg[j] = narg[j];
    }
    if (r + pstep < re) load_rows(r + pstep);
    /* uniform: every row of the step lies in [rs, re) */
    const bool full = r + pstep <= re;
#pragma unroll
    for (int sb = 0; sb < ROWS; sb += G) {
      qk_u32 kw[G], home[G], pay[G];
      QkBucket bk0[G], bk1[G];
      bool hit[G];
      if constexpr (PK) {
        /* kw: the expected slot word less the code for distance 0; home:
         * the home bucket's first slot */
#pragma unroll
        for (int j = 0; j < G; j++) {
          const qk_u32 h = ckey[sb + j];
          kw[j] = (0x40000u | (h & rmask)) << 13;
          home[j] = ((h >> rb) & (nbk - 1u)) << 2;
        }
#pragma unroll
        for (int j = 0; j < G; j++) {
          bk0[j] = qk_bucket_read(tk, home[j]);
          bk1[j] = qk_bucket_read(tk, (home[j] + QK_BUCKET) & smask);
        }
#pragma unroll
        for (int j = 0; j < G; j++) {
          const qk_u32 e0 = kw[j], e1 = e0 + dstep;
          qk_u32 d = min(min(bk0[j].w[0] - e0, bk0[j].w[1] - e0), min(bk0[j].w[2] - e0, bk0[j].w[3] - e0));
          d = min(d, min(min(bk1[j].w[0] - e1, bk1[j].w[1] - e1), min(bk1[j].w[2] - e1, bk1[j].w[3] - e1)));
          if (d >= 8192u) {
            /* not in the first two buckets: on while none of the buckets
             * read had an empty slot (rare) */
            qk_u32 z = min(min(min(bk0[j].w[0], bk0[j].w[1]), min(bk0[j].w[2], bk0[j].w[3])),
                           min(min(bk1[j].w[0], bk1[j].w[1]), min(bk1[j].w[2], bk1[j].w[3])));
            qk_u32 b = (home[j] + QK_BUCKET) & smask, e = e1;
            for (qk_u32 n = 2; z != 0u && d >= 8192u && n < maxd && n < nbk; n++) {
              b = (b + QK_BUCKET) & smask;
              e += dstep;
              const QkBucket nb = qk_bucket_read(tk, b);
              d = min(d, min(min(nb.w[0] - e, nb.w[1] - e), min(nb.w[2] - e, nb.w[3] - e)));
              z = min(min(nb.w[0], nb.w[1]), min(nb.w[2], nb.w[3]));
            }
          }
          hit[j] = d < range && (full || (r + (qk_u64)(sb + j) * WG + threadIdx.x) < re);
          pay[j] = d;
        }
      } else {
#pragma unroll
      for (int j = 0; j < G; j++) {
        const qk_u32 h = ckey[sb + j];
        kw[j] = (h & low) | 0x80000000u;
        home[j] = (h >> (32 - pbits - sbits)) & smask & ~(QK_BUCKET - 1u);
      }
      /* the home bucket and the next of each of the G rows; further
       * buckets follow only while both are full without the key (rare) */
#pragma unroll
      for (int j = 0; j < G; j++) {
        bk0[j] = qk_bucket_read(tk, home[j]);
        bk1[j] = qk_bucket_read(tk, (home[j] + QK_BUCKET) & smask);
      }
#pragma unroll
      for (int j = 0; j < G; j++) {
        const qk_u32 m0 = qk_bucket_match(bk0[j], kw[j]), m1 = qk_bucket_match(bk1[j], kw[j]);
        qk_u32 b = m0 ? home[j] : (home[j] + QK_BUCKET) & smask;
        qk_u32 m = m0 ? m0 : m1;
        if (!m) {
          qk_u32 e = qk_bucket_match(bk0[j], 0u) | qk_bucket_match(bk1[j], 0u);
          for (qk_u32 n = 2 * QK_BUCKET; !m && !e && n < S; n += QK_BUCKET) {
            b = (b + QK_BUCKET) & smask;
            QkBucket nb = qk_bucket_read(tk, b);
            m = qk_bucket_match(nb, kw[j]);
            e = qk_bucket_match(nb, 0u);
          }
        }
        qk_u32 idx = b + (qk_u32)__builtin_ctz(m | 0x100u);
        qk_u32 pj = m ? (per_row ? tp32[idx] : (qk_u32)tp16[idx]) : 0u;
        hit[j] = m && (full || (r + (qk_u64)(sb + j) * WG + threadIdx.x) < re) && pj < range;
        pay[j] = pj;
      }
      }
      if (pk_cs) {
        /* count << 40 | the argument, in one atomic */
#pragma unroll
        for (int j = 0; j < G; j++)
          if (hit[j]) atomicAdd(&my_sum[pay[j]], (1ULL << 40) | (qk_u64)carg[sb + j]);
        continue;
      }
#pragma unroll
      for (int j = 0; j < G; j++) {
        if (!hit[j]) continue;
        if (per_row) atomicAdd(&g_cnt[pay[j]], 1ULL);
        else atomicAdd(&my_cnt[pay[j]], 1u);
      }
      for (qk_u32 c = 0; c < nargs; c++) {
        qk_u32 at = (types >> (8 * c)) & 0xffu;
        qk_u32 v[G];
#pragma unroll
        for (int j = 0; j < G; j++) {
          qk_u64 x = r + (qk_u64)(sb + j) * WG + threadIdx.x;
          v[j] = c == 0 ? carg[sb + j] : rowsbuf[(x < re ? x : rs) * W + 1 + c];
        }
#pragma unroll
        for (int j = 0; j < G; j++) {
          if (!hit[j]) continue;
          qk_u32 g = pay[j];
          qk_u64 add = at == QK_T_I32 ? (qk_u64)(qk_s64)(qk_s32)v[j] : (qk_u64)v[j];
          qk_u32 code = qk_code32(v[j], at);
          if (per_row) {
            if (do_sum) atomicAdd(&g_sum[(qk_u64)c * range + g], add);
            if (do_mm) {
              atomicMin(&g_min[(qk_u64)c * range + g], code);
              atomicMax(&g_max[(qk_u64)c * range + g], code);
            }
          } else {
            if (do_sum) atomicAdd(&my_sum[c * range + g], add);
            if (do_mm) {
              atomicMin(&my_min[c * range + g], code);
              atomicMax(&my_max[c * range + g], code);
            }
          }
        }
      }
    }
  }
  }
  if (per_row) return;
  __syncthreads();
  if (pk_cs) {
    for (qk_u32 i = threadIdx.x; i < range; i += WG) {
      qk_u64 cnt = 0, sum = 0;
      for (qk_u32 k = 0; k < C; k++) {
        const qk_u64 x = l_sum[k * sstr + i];
        cnt += x >> 40;
        sum += x & ((1ULL << 40) - 1ULL);
      }
      if (!cnt) continue;
      atomicAdd(&g_cnt[i], cnt);
      atomicAdd(&g_sum[i], sum);
    }
    return;
  }
  for (qk_u32 i = threadIdx.x; i < range; i += WG) {
    qk_u32 cnt = 0;
    for (qk_u32 k = 0; k < C; k++) cnt += l_cnt[k * cstr + i];
    if (!cnt) continue;
    atomicAdd(&g_cnt[i], (qk_u64)cnt);
    for (qk_u32 c = 0; c < nargs; c++) {
      const qk_u32 w = c * range + i;
      if (do_sum) {
        qk_u64 sum = 0;
        for (qk_u32 k = 0; k < C; k++) sum += l_sum[k * sstr + w];
        atomicAdd(&g_sum[(qk_u64)c * range + i], sum);
      }
      if (do_mm) {
        qk_u32 mn = 0xffffffffu, mx = 0u;
        for (qk_u32 k = 0; k < C; k++) {
          mn = min(mn, l_min[k * mstr + w]);
          mx = max(mx, l_max[k * mstr + w]);
        }
        atomicMin(&g_min[(qk_u64)c * range + i], mn);
        atomicMax(&g_max[(qk_u64)c * range + i], mx);
      }
    }
  }
}

/* instances: generic, then group mode with 0-2 args (sums, min/max) and row
 * mode with 0-1 args; the host picks join_part_probe_<args><s|m|sm>[_row]
 * when one matches and the generic form otherwise.  Each comes as a
 * 256-thread block of 16 rows a thread per step, looked up 4 at a time,
 * and, suffixed _512, as a 512-thread block of 8 rows a thread looked up 2
 * at a time (the same LDS per block, fewer registers a thread, held to 80
 * so that three blocks, 6 waves a SIMD, fit a CU: twice the waves per CU). */
#define QK_PROBE_INSTANCE(name, na, sums, mm, row)                                                                  \
  extern "C" __global__ void LB k_join_part_probe##name(QK_Args a)                                                 \
  {                                                                                                                \
    qk_part_probe<na, sums, mm, row, QK_WG, 16, 4>(a);                                                                \
  }                                                                                                                \
  extern "C" __global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6)))                        \
      k_join_part_probe##name##_512(QK_Args a)                                                                     \
  {                                                                                                                \
    qk_part_probe<na, sums, mm, row, 512, 8, 2>(a);                                                                   \
  }
/* the packed-table forms of the group-mode shapes, 512-thread blocks */
#define QK_PROBE_PK_INSTANCE(name, na, sums, mm)                                                                     \
  extern "C" __global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6)))                        \
      k_join_part_probe_pk##name##_512(QK_Args a)                                                                  \
  {                                                                                                                \
    qk_part_probe<na, sums, mm, false, 512, 8, 2, true>(a);                                                           \
  }
QK_PROBE_PK_INSTANCE(_0, 0, false, false)
QK_PROBE_PK_INSTANCE(_1s, 1, true, false)
QK_PROBE_PK_INSTANCE(_1m, 1, false, true)
QK_PROBE_PK_INSTANCE(_1sm, 1, true, true)
QK_PROBE_PK_INSTANCE(_2s, 2, true, false)
QK_PROBE_PK_INSTANCE(_2sm, 2, true, true)
QK_PROBE_INSTANCE(, -1, false, false, false)
QK_PROBE_INSTANCE(_0, 0, false, false, false)
QK_PROBE_INSTANCE(_1s, 1, true, false, false)
QK_PROBE_INSTANCE(_1m, 1, false, true, false)
QK_PROBE_INSTANCE(_1sm, 1, true, true, false)
QK_PROBE_INSTANCE(_2s, 2, true, false, false)
QK_PROBE_INSTANCE(_2sm, 2, true, true, false)
QK_PROBE_INSTANCE(_0_row, 0, false, false, true)
QK_PROBE_INSTANCE(_1s_row, 1, true, false, true)
QK_PROBE_INSTANCE(_1sm_row, 1, true, true, true)

/* ------------------------------------------------------------------ */
/* hash GROUP BY (query_exec.c:10258-12300, general path: slot          */
/* assignment by CAS, then compaction and reduction)                    */
/* Keys are 64-bit codes: integers sign- or zero-extended, floats as    */
/* the bits of the value widened to f64 with -0.0 folded into 0.0.      */
/* Two levels: each block keeps an LDS table of QK_W_RANGE slots that   */
/* absorbs frequent keys; rows whose key finds no LDS slot within four  */
/* probes go to a global open-addressing table of 2^QK_W_LOG2 slots in  */
/* HBM.  The code ~0 (a 64-bit key only) has its own extra slot.        */
/* Global layout, C = 2^LOG2 + 1 slots: key u64[C] | min u64[E][C] |    */
/* kmin u64[T][C] | count u64[C] | sum u64[E][C] | max u64[E][C] |      */
/* kmax u64[T][C]; sums are exact 64-bit integers for U32/I32 args and  */
/* f64 bits otherwise; min/max are order-preserving codes of the f64    */
/* value.  T is 0 except for tuple keys (below).                        */
/* Several GROUP BY columns (binding 4: u32 words mode, K, then per key */
/* slot | type << 8, code minimum, range):                              */
/*  mode 1, digits: every key a U32/I32 column with values in [minimum, */
/*   minimum + range); the code is the keys' digits, key 0 most         */
/*   significant (the host keeps the product of the ranges below 2^64). */
/*  mode 2, tuple: any key types; the code is a 64-bit hash of the key  */
/*   codes, and each slot keeps the minimum and maximum of every key    */
/*   code (T = K), so the host sees whether two different keys ever     */
/*   shared a code (then it rehashes with another seed, word LOG2 >> 8).*/
/*   Tuple keys skip the LDS stage.                                     */
/* A digit outside its range (a stale zone map) sets flags[1].          */
/* ------------------------------------------------------------------ */
#define QK_MAX_KEYS 4
#define QK_HASH_EMPTY 0xffffffffffffffffULL
#define QK_HASH_MAX_ARGS 4

__device__ __forceinline__ qk_u64 qk_fmix64(qk_u64 k)
{
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

__device__ __forceinline__ qk_u64 qk_key_code_f64(double d) { return d == 0.0 ? 0ULL : (qk_u64)__double_as_longlong(d); }

/* key codes of the four rows of quad q */
__device__ __forceinline__ void qk_load4_code(const QK_Args& a, qk_u32 slot, qk_u32 type, qk_u64 q, qk_u64 c[4])
{
  const qk_u64 base = a.ptr[QK_B_COL0 + slot];
  if (type == QK_T_U32 || type == QK_T_I32 || type == QK_T_F32) {
    const qk_u32x4 x = qk_ld(qk_gptr<const qk_u32x4>(base) + q);
    const qk_u32 w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int k = 0; k < 4; k++)
      c[k] = type == QK_T_U32 ? (qk_u64)w[k]
                              : (type == QK_T_I32 ? (qk_u64)(qk_s64)(qk_s32)w[k] : qk_key_code_f64((double)__uint_as_float(w[k])));
  } else {
    const qk_u64x2* p = qk_gptr<const qk_u64x2>(base) + 2 * q;
    const qk_u64x2 x0 = qk_ld(p), x1 = qk_ld(p + 1);
    const qk_u64 w[4] = {x0.x, x0.y, x1.x, x1.y};
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = type == QK_T_F64 ? qk_key_code_f64(__longlong_as_double((qk_s64)w[k])) : w[k];
  }
}

__device__ __forceinline__ qk_u64 qk_load1_code(const QK_Args& a, qk_u32 slot, qk_u32 type, qk_u64 row)
{
  const qk_u64 base = a.ptr[QK_B_COL0 + slot];
  switch (type) {
    case QK_T_U32: return (qk_u64)qk_gptr<const qk_u32>(base)[row];
    case QK_T_I32: return (qk_u64)(qk_s64)qk_gptr<const qk_s32>(base)[row];
    case QK_T_F32: return qk_key_code_f64((double)qk_gptr<const float>(base)[row]);
    case QK_T_F64: return qk_key_code_f64(qk_gptr<const double>(base)[row]);
    default: return qk_gptr<const qk_u64>(base)[row];
  }
}

/* the global slot of `key`, inserting it; ~0 when the table is full.  Slots
 * change only from EMPTY to a key, so a stale read can only show EMPTY, and
 * the CAS then returns the key that is really there. */
__device__ __forceinline__ qk_u64 qk_hash_slot(qk_u64* gkey, qk_u32 log2, qk_u64 key, qk_u32* flags)
{
  const qk_u64 cap = 1ULL << log2, mask = cap - 1;
  if (key == QK_HASH_EMPTY) return cap;
  qk_u64 s = qk_fmix64(key) & mask;
  for (qk_u64 n = 0; n < cap; n++) {
    qk_u64 cur = gkey[s];
    if (cur == key) return s;
    if (cur == QK_HASH_EMPTY) {
      qk_u64 prev = atomicCAS((unsigned long long*)&gkey[s], QK_HASH_EMPTY, key);
      if (prev == QK_HASH_EMPTY || prev == key) return s;
    }
    s = (s + 1) & mask;
  }
  atomicOr(&flags[0], 1u);
  return ~0ULL;
}

struct QkHashCtx {
  qk_u32 E, types, flags, L, log2, T;
  qk_u64 C;
  qk_u64* gkey;
  qk_u64* gmin;
  qk_u64* gkmin;
  qk_u64* gcnt;
  qk_u64* gsum;
  qk_u64* gmax;
  qk_u64* gkmax;
  qk_u32* gflags;
  qk_u64* lkey;
  qk_u32* lcnt;
  qk_u64* lsum;
  qk_u64* lmin;
  qk_u64* lmax;
};

__device__ __forceinline__ void qk_hash_global_add(const QkHashCtx& h, qk_u64 key, qk_u64 cnt, const qk_u64* sum,
                                                   const qk_u64* mn, const qk_u64* mx, const qk_u64* kc = nullptr)
{
  const qk_u64 s = qk_hash_slot(h.gkey, h.log2, key, h.gflags);
  if (s == ~0ULL) return;
  atomicAdd((unsigned long long*)&h.gcnt[s], (unsigned long long)cnt);
  if (kc) {
    for (qk_u32 i = 0; i < h.T; i++) {
      atomicMin((unsigned long long*)&h.gkmin[(qk_u64)i * h.C + s], (unsigned long long)kc[i]);
      atomicMax((unsigned long long*)&h.gkmax[(qk_u64)i * h.C + s], (unsigned long long)kc[i]);
    }
  }
  for (qk_u32 c = 0; c < h.E; c++) {
    const qk_u32 at = (h.types >> (8 * c)) & 0xffu;
    if (h.flags & 1u) {
      if (qk_type_is_int32(at)) atomicAdd((unsigned long long*)&h.gsum[(qk_u64)c * h.C + s], (unsigned long long)sum[c]);
      else unsafeAtomicAdd((double*)&h.gsum[(qk_u64)c * h.C + s], __longlong_as_double((qk_s64)sum[c]));
    }
    if (h.flags & 2u) {
      atomicMin((unsigned long long*)&h.gmin[(qk_u64)c * h.C + s], (unsigned long long)mn[c]);
      atomicMax((unsigned long long*)&h.gmax[(qk_u64)c * h.C + s], (unsigned long long)mx[c]);
    }
  }
}

/* one row: into the block's LDS table when its key has (or gets) a slot
 * there within four probes, else straight into the global table */
__device__ __forceinline__ void qk_hash_row(const QkHashCtx& h, qk_u64 key, const double* v, const qk_u64* kc = nullptr)
{
  qk_u64 sum[QK_HASH_MAX_ARGS], code[QK_HASH_MAX_ARGS];
#pragma unroll
  for (int c = 0; c < QK_HASH_MAX_ARGS; c++) {
    const qk_u32 at = (h.types >> (8 * c)) & 0xffu;
    sum[c] = qk_type_is_int32(at) ? (qk_u64)(qk_s64)v[c] : (qk_u64)__double_as_longlong(v[c]);
    code[c] = qk_f64_key(v[c]);
  }
  if (h.L && key != QK_HASH_EMPTY) {
    const qk_u32 lmask = h.L - 1u;
    qk_u32 s = (qk_u32)qk_fmix64(key) & lmask;
#pragma unroll 1
    for (int p = 0; p < 4; p++, s = (s + 1u) & lmask) {
      qk_u64 cur = h.lkey[s];
      if (cur == QK_HASH_EMPTY) cur = atomicCAS((unsigned long long*)&h.lkey[s], QK_HASH_EMPTY, key);
      if (cur != QK_HASH_EMPTY && cur != key) continue;
      atomicAdd(&h.lcnt[s], 1u);
      for (qk_u32 c = 0; c < h.E; c++) {
        const qk_u32 at = (h.types >> (8 * c)) & 0xffu;
        if (h.flags & 1u) {
          if (qk_type_is_int32(at)) atomicAdd((unsigned long long*)&h.lsum[c * h.L + s], (unsigned long long)sum[c]);
          else atomicAdd((double*)&h.lsum[c * h.L + s], v[c]);
        }
        if (h.flags & 2u) {
          atomicMin((unsigned long long*)&h.lmin[c * h.L + s], (unsigned long long)code[c]);
          atomicMax((unsigned long long*)&h.lmax[c * h.L + s], (unsigned long long)code[c]);
        }
      }
      return;
    }
  }
  qk_hash_global_add(h, key, 1, sum, code, code, kc);
}

/* the group code of several keys (binding 4's descriptor) for one row:
 * digits or the tuple hash (with the key codes in kc); *ok cleared when a
 * digit is outside its range */
__device__ __forceinline__ qk_u64 qk_multi_code1(const QK_Args& a, const qk_u32* kd, qk_u64 row, qk_u64* kc, bool* ok)
{
  const qk_u32 mode = kd[0], nk = kd[1];
  qk_u64 code = mode == 2 ? (qk_u64)kd[2 + 3 * nk] * 0x9E3779B97F4A7C15ULL : 0ULL;
  for (qk_u32 i = 0; i < nk; i++) {
    const qk_u32 d = kd[2 + 3 * i], t = (d >> 8) & 0xfu;
    const qk_u64 c = qk_load1_code(a, d & 0xffu, t, row);
    if (mode == 1) {
      const qk_u32 dg = qk_code32((qk_u32)c, t) - kd[3 + 3 * i], r = kd[4 + 3 * i];
      if (dg >= r) *ok = false;
      code = code * r + dg;
    } else {
      kc[i] = c;
      code = qk_fmix64(code ^ (c + 0x9E3779B97F4A7C15ULL * (i + 1)));
    }
  }
  return code;
}

/* the same for the four rows of quad q: codes c[4], key codes kc[key][4];
 * a row with a digit outside its range leaves mask m */
__device__ __forceinline__ void qk_multi_code4(const QK_Args& a, const qk_u32* kd, qk_u64 q, qk_u64 c[4],
                                               qk_u64 kc[QK_MAX_KEYS][4], qk_u32& m, qk_u32& bad)
{
  const qk_u32 mode = kd[0], nk = kd[1];
  const qk_u64 seed = mode == 2 ? (qk_u64)kd[2 + 3 * nk] * 0x9E3779B97F4A7C15ULL : 0ULL;
#pragma unroll
  for (int k = 0; k < 4; k++) c[k] = seed;
  for (qk_u32 i = 0; i < nk && i < QK_MAX_KEYS; i++) {
    const qk_u32 d = kd[2 + 3 * i], t = (d >> 8) & 0xfu;
    qk_u64 x[4];
    qk_load4_code(a, d & 0xffu, t, q, x);
    if (mode == 1) {
      const qk_u32 mn = kd[3 + 3 * i], r = kd[4 + 3 * i];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const qk_u32 dg = qk_code32((qk_u32)x[k], t) - mn;
        if (dg >= r && (m & (1u << k))) {
          m &= ~(1u << k);
          bad = 1u;
        }
        c[k] = c[k] * r + dg;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        kc[i][k] = x[k];
        c[k] = qk_fmix64(c[k] ^ (x[k] + 0x9E3779B97F4A7C15ULL * (i + 1)));
      }
    }
  }
}

/* aggregate_hash: binding 0 leaves, 1 table (layout above), 2 flags
 * (u32[0]: global table full), 3 row list (u64 ids, optional: then push 0
 * is the list length), 5 skip bits, 6.. columns; push 0 rows, 3/4 skip,
 * 5 first row, 7 leaf count; words ARGS (count | slots << 8), TYPES, KEY
 * (key slot | key type << 8 | flags << 16: 1 sums, 2 min/max, 4 no key,
 * every row in group 0), LOG2, RANGE (LDS slots, a power of two or 0).
 * LDS: key u64[L] | min u64[E][L] | max u64[E][L] | sum u64[E][L] |
 * count u32[L] */
template <bool TERMS>
__device__ __forceinline__ void qk_aggregate_hash(const QK_Args& a)
{
  extern __shared__ qk_u64 lds_u64[];
  QkHashCtx h;
  h.E = a.words[QK_W_ARGS] & 0xffu;
  if (h.E > QK_HASH_MAX_ARGS) h.E = QK_HASH_MAX_ARGS;
  const qk_u32 slots = a.words[QK_W_ARGS] >> 8;
  h.types = a.words[QK_W_TYPES];
  const qk_u32 key_slot = a.words[QK_W_KEY] & 0xffu;
  /* type bytes: column type, encoding << 4 (encoded columns are codes, U32) */
  const qk_u32 key_type = (a.words[QK_W_KEY] >> 8) & 0xfu;
  const qk_u32 key_enc = (a.words[QK_W_KEY] >> 12) & 3u;
  h.flags = a.words[QK_W_KEY] >> 16;
  h.log2 = a.words[QK_W_LOG2] & 0xffu;
  h.L = a.words[QK_W_RANGE];
  const qk_u32* kd = a.ptr[4] ? qk_ptr<qk_u32>(a, 4) : nullptr;
  const bool tuple = kd && kd[0] == 2u;
  h.T = tuple ? kd[1] : 0u;
  if (tuple) h.L = 0;
  h.C = (1ULL << h.log2) + 1;
  h.gkey = qk_wptr<qk_u64>(a, 1);
  h.gmin = h.gkey + h.C;
  h.gkmin = h.gmin + (qk_u64)h.E * h.C;
  h.gcnt = h.gkmin + (qk_u64)h.T * h.C;
  h.gsum = h.gcnt + h.C;
  h.gmax = h.gsum + (qk_u64)h.E * h.C;
  h.gkmax = h.gmax + (qk_u64)h.E * h.C;
  h.gflags = qk_wptr<qk_u32>(a, 2);
  h.lkey = lds_u64;
  h.lmin = h.lkey + h.L;
  h.lmax = h.lmin + (qk_u64)h.E * h.L;
  h.lsum = h.lmax + (qk_u64)h.E * h.L;
  h.lcnt = (qk_u32*)(h.lsum + (qk_u64)h.E * h.L);
  for (qk_u32 i = threadIdx.x; i < h.L; i += QK_WG) {
    h.lkey[i] = QK_HASH_EMPTY;
    h.lcnt[i] = 0;
    for (qk_u32 c = 0; c < h.E; c++) {
      h.lsum[c * h.L + i] = 0;
      h.lmin[c * h.L + i] = QK_HASH_EMPTY;
      h.lmax[c * h.L + i] = 0;
    }
  }
  __syncthreads();
  const bool no_key = (h.flags & 4u) != 0;
  const qk_u64* list = a.ptr[3] ? qk_ptr<qk_u64>(a, 3) : nullptr;
  if (list) {
    const qk_u64 n = qk_push(a, QK_P_ROWS);
    for (qk_u64 i = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; i < n; i += (qk_u64)gridDim.x * QK_WG) {
      const qk_u64 row = list[i];
      double v[QK_HASH_MAX_ARGS] = {0.0, 0.0, 0.0, 0.0};
      for (qk_u32 c = 0; c < h.E; c++)
        v[c] = qk_load1_f64(a, (slots >> (4 * c)) & 0xfu, (h.types >> (8 * c)) & 0xffu, row);
      if (kd) {
        qk_u64 kc[QK_MAX_KEYS];
        bool ok = true;
        const qk_u64 code = qk_multi_code1(a, kd, row, kc, &ok);
        if (ok) qk_hash_row(h, code, v, tuple ? kc : nullptr);
        else atomicOr(&h.gflags[1], 1u);
      } else {
        qk_hash_row(h, no_key ? 0ULL : qk_load1_code(a, key_slot, key_type, row), v);
      }
    }
  } else {
    const qk_u64 rows = qk_push(a, QK_P_ROWS);
    const qk_u64 r0 = qk_push(a, QK_P_ROW_BEGIN);
    const qk_u64 r1 = r0 + rows;
    const QkLeafSet L = qk_load_leaves<TERMS>(qk_ptr<qk_u32>(a, QK_B_LEAVES), (qk_u32)qk_push(a, QK_P_LEAVES));
    qk_u64 b0, b1;
    qk_block_range(r0 / 4, (r1 + 3) / 4, b0, b1);
    for (qk_u64 s = b0; s < b1; s += (qk_u64)QK_WG * QK_UNROLL) {
      qk_u64 q[QK_UNROLL];
      qk_u32 m[QK_UNROLL];
      qk_step_quads(a, s, b1, r0, r1, q, m);
      if (L.n) qk_leaves_step<TERMS>(a, L, q, m);
#pragma unroll 1
      for (int j = 0; j < QK_UNROLL; j++) {
        if (!m[j]) continue;
        qk_u64 kc[4] = {0, 0, 0, 0};
        qk_u64 kcs[QK_MAX_KEYS][4];
        if (kd) {
          qk_u32 bad = 0u;
          qk_multi_code4(a, kd, q[j], kc, kcs, m[j], bad);
          if (bad) atomicOr(&h.gflags[1], 1u);
        } else if (!no_key) {
          qk_load4_code(a, key_slot, key_type, q[j], kc);
        }
        double v[QK_HASH_MAX_ARGS][4];
#pragma unroll
        for (int c = 0; c < QK_HASH_MAX_ARGS; c++) {
          if ((qk_u32)c < h.E) qk_load4_f64(a, (slots >> (4 * c)) & 0xfu, (h.types >> (8 * c)) & 0xffu, q[j], v[c]);
          else v[c][0] = v[c][1] = v[c][2] = v[c][3] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
          if (!(m[j] & (1u << k))) continue;
          const double vk[QK_HASH_MAX_ARGS] = {v[0][k], v[1][k], v[2][k], v[3][k]};
          if (tuple) {
            qk_u64 kk[QK_MAX_KEYS];
#pragma unroll
            for (int i = 0; i < QK_MAX_KEYS; i++) kk[i] = kcs[i][k];
            qk_hash_row(h, kc[k], vk, kk);
          } else {
            qk_hash_row(h, kc[k], vk);
          }
        }
      }
    }
  }
  /* flush the LDS table into the global one */
  __syncthreads();
  for (qk_u32 i = threadIdx.x; i < h.L; i += QK_WG) {
    const qk_u32 cnt = h.lcnt[i];
    if (!cnt) continue;
    qk_u64 sum[QK_HASH_MAX_ARGS], mn[QK_HASH_MAX_ARGS], mx[QK_HASH_MAX_ARGS];
    for (qk_u32 c = 0; c < QK_HASH_MAX_ARGS; c++) {
      const bool in = c < h.E;
      sum[c] = in ? h.lsum[c * h.L + i] : 0;
      mn[c] = in ? h.lmin[c * h.L + i] : QK_HASH_EMPTY;
      mx[c] = in ? h.lmax[c * h.L + i] : 0;
    }
    qk_hash_global_add(h, h.lkey[i], cnt, sum, mn, mx);
  }
}

extern "C" __global__ void LB k_aggregate_hash(QK_Args a) { qk_aggregate_hash<false>(a); }
/* the same over a WHERE of several conjunctions */
extern "C" __global__ void LB k_aggregate_hash_or(QK_Args a) { qk_aggregate_hash<true>(a); }

/* hash_compact: the occupied slots of aggregate_hash's table, in slot
 * order.  binding 1 table, 2 block counts (u64 per block), 3 out records
 * (u64[2 + 3E + 2T] each: key, count, sums, mins, maxs, key minima, key
 * maxima); push 6 mode (1 block counts, 2 write), words ARGS (count | T
 * << 8), LOG2.  A slot is occupied when its count is nonzero (rows only
 * ever add). */
extern "C" __global__ void LB k_hash_compact(QK_Args a)
{
  __shared__ qk_u32 s_scan[QK_WG / 64];
  __shared__ qk_u64 s_red[QK_WG / 64];
  const qk_u32 E = a.words[QK_W_ARGS] & 0xffu;
  const qk_u32 T = (a.words[QK_W_ARGS] >> 8) & 0xffu;
  const qk_u32 log2 = a.words[QK_W_LOG2] & 0xffu;
  const qk_u64 C = (1ULL << log2) + 1;
  const qk_u64* gkey = qk_ptr<qk_u64>(a, 1);
  const qk_u64* gmin = gkey + C;
  const qk_u64* gkmin = gmin + (qk_u64)E * C;
  const qk_u64* gcnt = gkmin + (qk_u64)T * C;
  const qk_u64* gsum = gcnt + C;
  const qk_u64* gmax = gsum + (qk_u64)E * C;
  const qk_u64* gkmax = gmax + (qk_u64)E * C;
  qk_u64* bc = qk_wptr<qk_u64>(a, 2);
  const qk_u64 mode = qk_push(a, QK_P_MODE);
  const qk_u64 per = (C + gridDim.x - 1) / gridDim.x;
  const qk_u64 s0 = per * blockIdx.x < C ? per * blockIdx.x : C;
  const qk_u64 s1 = s0 + per < C ? s0 + per : C;
  if (mode == 1) {
    qk_u64 n = 0;
    for (qk_u64 s = s0 + threadIdx.x; s < s1; s += QK_WG) n += gcnt[s] ? 1 : 0;
    n = qk_block_sum_u64(n, s_red);
    if (threadIdx.x == 0) bc[blockIdx.x] = n;
    return;
  }
  qk_u64 base = 0;
  for (qk_u32 b = threadIdx.x; b < blockIdx.x; b += QK_WG) base += bc[b];
  base = qk_block_sum_u64(base, s_red);
  qk_u64* out = qk_wptr<qk_u64>(a, 3);
  const qk_u64 rw = 2 + 3 * (qk_u64)E + 2 * (qk_u64)T;
  const qk_u64 cap = a.bytes[3] / (8ULL * rw);
  for (qk_u64 c0 = s0; c0 < s1; c0 += QK_WG) {
    const qk_u64 s = c0 + threadIdx.x;
    const qk_u32 occ = (s < s1 && gcnt[s]) ? 1u : 0u;
    qk_u32 total = 0;
    const qk_u32 pre = qk_block_excl_scan_u32(occ, s_scan, total);
    if (occ) {
      const qk_u64 g = base + pre;
      if (g < cap) {
        qk_u64* r = out + g * rw;
        r[0] = gkey[s];
        r[1] = gcnt[s];
        for (qk_u32 c = 0; c < E; c++) {
          r[2 + c] = gsum[(qk_u64)c * C + s];
          r[2 + E + c] = gmin[(qk_u64)c * C + s];
          r[2 + 2 * E + c] = gmax[(qk_u64)c * C + s];
        }
        for (qk_u32 i = 0; i < T; i++) {
          r[2 + 3 * E + i] = gkmin[(qk_u64)i * C + s];
          r[2 + 3 * E + T + i] = gkmax[(qk_u64)i * C + s];
        }
      }
    }
    base += total;
  }
}

/* ------------------------------------------------------------------ */
/* shared scan: many aggregate queries over the same resident columns  */
/* in one launch (the reference fuses one WHERE with one aggregate per */
/* kernel, query_kernel.c:1433-1600; here one pass serves a batch).    */
/* Program (u32 words, binding 0):                                      */
/*   [0] queries Q  [1] accumulator words W  [2] offset of the op table */
/*   [3] staged columns U  [4] bytes of one stage buffer                */
/*   [5] LDS accumulator words WL (W with every replica)                */
/*   [6] offset of the chunk table, [7] its chunks C: a tile's stage as  */
/*   1 KB chunks, entry f = staged column u | chunk k << 16 (bytes       */
/*   k * 1024 on of the column's rows in the tile), or 0x8000 | j << 16  */
/*   for chunk j of the null bitmaps                                    */
/*   [8] null bitmaps NB staged (of the nullable columns WHERE leaves    */
/*   read), [9] offset of their table (per bitmap: table column, byte   */
/*   offset of its bitmap / 16), [10] byte offset of the bitmap area in */
/*   a stage buffer, [11] its 1 KB chunks: a tile of T rows holds bitmap */
/*   s as T / 8 bytes at area + s * T / 8 (bit r of word r / 32 set =   */
/*   tile row r NULL)                                                   */
/*   [12 + q] offset of query q's record                                */
/*   [12 + Q + 3u] staged column u: table column, width in bytes (4 or  */
/*   8), byte offset of its rows in a stage buffer (16-byte aligned)    */
/* query record: [0] leaves L (<= 8) | terms T << 8 (0 = the AND of   */
/*   the leaves, else the rows of any term whose leaves all pass),      */
/*   [1] flags (1 grouped, 2 sums,                                      */
/*   4 min/max, 8 grouped rows walked compactly where a wave's rows are */
/*   sparse: a lane visits only the rows that passed), [2] key (byte   */
/*   offset of its rows in a stage buffer), [3] key type,               */
/*   [4] key_min code,                                                  */
/*   [5] groups R (1 when ungrouped), [6] args A (<= 2), [7] arg 0      */
/*   (stage byte offset), [8] arg type 0, [9] arg 1, [10] arg type 1,   */
/*   [11] offset of its words among the W, [12] offset of its LDS      */
/*   words, [13] replicas (a power of two; lane l of the owning wave    */
/*   adds into replica l mod replicas, the block folds them at its end),*/
/*   [14..] leaves, 6 words each (qe_simple_leaf_words layout, word 0  */
/*   the column's stage byte offset; word 5 of a nullable column's leaf */
/*   1 + its staged bitmap's byte offset in a stage buffer / 16),       */
/*   then the T terms' leaf bits, a byte each.                          */
/* Columns: binding 2 holds the base address of every table column.     */
/* A block walks tiles of 256 x ITEMS rows.  The rows of every staged   */
/* column of a tile are copied straight from HBM into an LDS stage      */
/* buffer (global_load_lds, 16 bytes per lane, no registers in          */
/* between), the copies of the next tile are issued into the other      */
/* buffer before the current one is evaluated, and the queries are then */
/* evaluated on the tile from LDS, each wave taking the next query of  */
/* the list (most expensive first) from an LDS counter until none is    */
/* left, over all the tile's rows, so each column is read from HBM once per */
/* launch however many queries read it and each query's record and     */
/* reductions are paid once per tile and wave.                          */
/* Accumulators of a query, u64 words from its offset: count[R], then   */
/* per arg sum[R], min[R], max[R] (sums exact integers for U32/I32 args */
/* and f64 bits otherwise; min/max order-preserving codes of the f64    */
/* value).  op table: one u32 per accumulator word (0 integer add,      */
/* 1 f64 add, 2 min, 3 max).  Each block writes its W words to          */
/* binding 1 [block][W]; shared_reduce folds them in block order.       */
/* LDS: accumulators u64[WL] (padded to 16 bytes) | stage buffer 0 |    */
/* stage buffer 1 (when double-buffered, push slot 6).                  */
/* ------------------------------------------------------------------ */
#define QK_SS_MAX_ARGS 2
#define QK_SS_HDR 12
#define QK_SS_REC 14

__device__ __forceinline__ bool qk_type_is_wide(qk_u32 t) { return t == QK_T_F64 || t == QK_T_I64 || t == QK_T_U64; }

/* wave64 reductions on DPP: xor-1 and xor-2 quad permutes, then the
 * half-row and row mirrors leave each 16-lane row's total in all of its
 * lanes, and four readlanes add the rows.  Every lane of the wave must be
 * active (callers are in uniform control flow). */
template <int CTRL>
__device__ __forceinline__ qk_u64 qk_dpp64(qk_u64 v)
{
  const qk_u32 lo = (qk_u32)__builtin_amdgcn_update_dpp(0, (int)(qk_u32)v, CTRL, 0xf, 0xf, false);
  const qk_u32 hi = (qk_u32)__builtin_amdgcn_update_dpp(0, (int)(qk_u32)(v >> 32), CTRL, 0xf, 0xf, false);
  return ((qk_u64)hi << 32) | lo;
}

__device__ __forceinline__ qk_u64 qk_readlane64(qk_u64 v, int lane)
{
  const qk_u32 lo = (qk_u32)__builtin_amdgcn_readlane((int)(qk_u32)v, lane);
  const qk_u32 hi = (qk_u32)__builtin_amdgcn_readlane((int)(qk_u32)(v >> 32), lane);
  return ((qk_u64)hi << 32) | lo;
}

#define QK_DPP_XOR1 0xB1
#define QK_DPP_XOR2 0x4E
#define QK_DPP_HALF_MIRROR 0x141
#define QK_DPP_MIRROR 0x140

__device__ __forceinline__ qk_u64 qk_wave_add_u64(qk_u64 v)
{
  v += qk_dpp64<QK_DPP_XOR1>(v);
  v += qk_dpp64<QK_DPP_XOR2>(v);
  v += qk_dpp64<QK_DPP_HALF_MIRROR>(v);
  v += qk_dpp64<QK_DPP_MIRROR>(v);
  return qk_readlane64(v, 0) + qk_readlane64(v, 16) + qk_readlane64(v, 32) + qk_readlane64(v, 48);
}

__device__ __forceinline__ double qk_wave_add_f64(double d)
{
  qk_u64 v = (qk_u64)__double_as_longlong(d);
  d += __longlong_as_double((qk_s64)qk_dpp64<QK_DPP_XOR1>(v));
  v = (qk_u64)__double_as_longlong(d);
  d += __longlong_as_double((qk_s64)qk_dpp64<QK_DPP_XOR2>(v));
  v = (qk_u64)__double_as_longlong(d);
  d += __longlong_as_double((qk_s64)qk_dpp64<QK_DPP_HALF_MIRROR>(v));
  v = (qk_u64)__double_as_longlong(d);
  d += __longlong_as_double((qk_s64)qk_dpp64<QK_DPP_MIRROR>(v));
  v = (qk_u64)__double_as_longlong(d);
  return (__longlong_as_double((qk_s64)qk_readlane64(v, 0)) + __longlong_as_double((qk_s64)qk_readlane64(v, 16))) +
         (__longlong_as_double((qk_s64)qk_readlane64(v, 32)) + __longlong_as_double((qk_s64)qk_readlane64(v, 48)));
}

/* 32-bit forms: one DPP-fed operation per step */
__device__ __forceinline__ qk_u32 qk_wave_add_u32(qk_u32 v)
{
  v += (qk_u32)__builtin_amdgcn_update_dpp(0, (int)v, QK_DPP_XOR1, 0xf, 0xf, false);
  v += (qk_u32)__builtin_amdgcn_update_dpp(0, (int)v, QK_DPP_XOR2, 0xf, 0xf, false);
  v += (qk_u32)__builtin_amdgcn_update_dpp(0, (int)v, QK_DPP_HALF_MIRROR, 0xf, 0xf, false);
  v += (qk_u32)__builtin_amdgcn_update_dpp(0, (int)v, QK_DPP_MIRROR, 0xf, 0xf, false);
  return (qk_u32)__builtin_amdgcn_readlane((int)v, 0) + (qk_u32)__builtin_amdgcn_readlane((int)v, 16) +
         (qk_u32)__builtin_amdgcn_readlane((int)v, 32) + (qk_u32)__builtin_amdgcn_readlane((int)v, 48);
}

__device__ __forceinline__ void qk_wave_minmax_u32(qk_u32& mn, qk_u32& mx)
{
#define QK_MM32_STEP(C)                                                                   \
  {                                                                                       \
    const qk_u32 a2 = (qk_u32)__builtin_amdgcn_update_dpp(0, (int)mn, C, 0xf, 0xf, false); \
    const qk_u32 b2 = (qk_u32)__builtin_amdgcn_update_dpp(0, (int)mx, C, 0xf, 0xf, false); \
    mn = a2 < mn ? a2 : mn;                                                               \
    mx = b2 > mx ? b2 : mx;                                                               \
  }
  QK_MM32_STEP(QK_DPP_XOR1)
  QK_MM32_STEP(QK_DPP_XOR2)
  QK_MM32_STEP(QK_DPP_HALF_MIRROR)
  QK_MM32_STEP(QK_DPP_MIRROR)
#undef QK_MM32_STEP
  qk_u32 n0 = (qk_u32)__builtin_amdgcn_readlane((int)mn, 0), n1 = (qk_u32)__builtin_amdgcn_readlane((int)mn, 16);
  qk_u32 n2 = (qk_u32)__builtin_amdgcn_readlane((int)mn, 32), n3 = (qk_u32)__builtin_amdgcn_readlane((int)mn, 48);
  qk_u32 x0 = (qk_u32)__builtin_amdgcn_readlane((int)mx, 0), x1 = (qk_u32)__builtin_amdgcn_readlane((int)mx, 16);
  qk_u32 x2 = (qk_u32)__builtin_amdgcn_readlane((int)mx, 32), x3 = (qk_u32)__builtin_amdgcn_readlane((int)mx, 48);
  n0 = n1 < n0 ? n1 : n0;
  n2 = n3 < n2 ? n3 : n2;
  mn = n2 < n0 ? n2 : n0;
  x0 = x1 > x0 ? x1 : x0;
  x2 = x3 > x2 ? x3 : x2;
  mx = x2 > x0 ? x2 : x0;
}

__device__ __forceinline__ void qk_wave_minmax_u64(qk_u64& mn, qk_u64& mx)
{
#define QK_MM_STEP(C)                            \
  {                                              \
    const qk_u64 a2 = qk_dpp64<C>(mn);           \
    const qk_u64 b2 = qk_dpp64<C>(mx);           \
    mn = a2 < mn ? a2 : mn;                      \
    mx = b2 > mx ? b2 : mx;                      \
  }
  QK_MM_STEP(QK_DPP_XOR1)
  QK_MM_STEP(QK_DPP_XOR2)
  QK_MM_STEP(QK_DPP_HALF_MIRROR)
  QK_MM_STEP(QK_DPP_MIRROR)
#undef QK_MM_STEP
  qk_u64 n0 = qk_readlane64(mn, 0), n1 = qk_readlane64(mn, 16), n2 = qk_readlane64(mn, 32), n3 = qk_readlane64(mn, 48);
  qk_u64 x0 = qk_readlane64(mx, 0), x1 = qk_readlane64(mx, 16), x2 = qk_readlane64(mx, 32), x3 = qk_readlane64(mx, 48);
  n0 = n1 < n0 ? n1 : n0;
  n2 = n3 < n2 ? n3 : n2;
  mn = n2 < n0 ? n2 : n0;
  x0 = x1 > x0 ? x1 : x0;
  x2 = x3 > x2 ? x3 : x2;
  mx = x2 > x0 ? x2 : x0;
}

/* value i of a staged column (its bytes at `p`, type `t`) as f64 */
__device__ __forceinline__ double qk_ss_f64(const unsigned char* p, qk_u32 t, qk_u32 i)
{
  if (qk_type_is_wide(t)) {
    const qk_u64 x = ((const qk_u64*)p)[i];
    return t == QK_T_F64 ? __longlong_as_double((qk_s64)x) : (t == QK_T_I64 ? (double)(qk_s64)x : (double)x);
  }
  const qk_u32 x = ((const qk_u32*)p)[i];
  return t == QK_T_U32 ? (double)x : (t == QK_T_I32 ? (double)(qk_s32)x : (double)__uint_as_float(x));
}

/* N of the lane's rows, i0 + i for i < N (row (i0 + i) * 64 + lane of the
 * tile), of a staged column as f64, with the column type dispatched once
 * for all of them.  Callers walk a lane's rows in groups of N <= 8 with i0
 * an unrolled constant, so at most 8 values are live at a time. */
template <int N, typename T>
__device__ __forceinline__ void qk_ss_rows_t(const unsigned char* p, qk_u32 lane, int i0, double* v)
{
  const T* sv = (const T*)p;
#pragma unroll
  for (int i = 0; i < N; i++) v[i] = (double)sv[(i0 + i) * 64 + lane];
}

template <int N>
__device__ __forceinline__ void qk_ss_rows(const unsigned char* p, qk_u32 t, qk_u32 lane, int i0, double* v)
{
  switch (t) {
    case QK_T_U32: qk_ss_rows_t<N, qk_u32>(p, lane, i0, v); break;
    case QK_T_I32: qk_ss_rows_t<N, qk_s32>(p, lane, i0, v); break;
    case QK_T_F32: qk_ss_rows_t<N, float>(p, lane, i0, v); break;
    case QK_T_F64: qk_ss_rows_t<N, double>(p, lane, i0, v); break;
    case QK_T_I64: qk_ss_rows_t<N, qk_s64>(p, lane, i0, v); break;
    default: qk_ss_rows_t<N, qk_u64>(p, lane, i0, v); break;
  }
}

/* clear the bits (i0 + i) of rows whose value fails `x[i] op tv` (op
 * dispatched once) */
template <int N>
__device__ __forceinline__ qk_u32 qk_ss_cmp(const double* x, qk_u32 op, double tv, qk_u32 m, int i0)
{
  switch (op) {
    case 0:
#pragma unroll
      for (int i = 0; i < N; i++) m &= (x[i] == tv) ? ~0u : ~(1u << (i0 + i));
      break;
    case 1:
#pragma unroll
      for (int i = 0; i < N; i++) m &= (x[i] < tv) ? ~0u : ~(1u << (i0 + i));
      break;
    case 2:
#pragma unroll
      for (int i = 0; i < N; i++) m &= (x[i] <= tv) ? ~0u : ~(1u << (i0 + i));
      break;
    case 3:
#pragma unroll
      for (int i = 0; i < N; i++) m &= (x[i] > tv) ? ~0u : ~(1u << (i0 + i));
      break;
    default:
#pragma unroll
      for (int i = 0; i < N; i++) m &= (x[i] >= tv) ? ~0u : ~(1u << (i0 + i));
      break;
  }
  return m;
}

/* the program and the column table do not change during a launch: read
 * through the constant address space, whose uniform loads the compiler
 * makes scalar loads */
#define QK_CONST __attribute__((address_space(4)))
typedef const QK_CONST qk_u32* qk_cu32p;
typedef const QK_CONST qk_u64* qk_cu64p;

/* issue the copies of tile `t` (rows base..base+tile) of every staged
 * column into the stage buffer at `buf`: 1 KB per wave instruction, the
 * chunks dealt round-robin over the block's waves; lanes past the
 * column's end read its last 16 bytes instead (those rows are masked) */
/* the address lane `lane` copies 16 bytes from for chunk j of a tile's
 * staged null bitmaps (tile rows from `base`): bitmap j * (64 / lpb) +
 * lane / lpb, bytes (lane mod lpb) * 16 on of the tile's part, lpb = tile
 * / 128 lanes a bitmap.  Lanes past the last bitmap read the last one
 * again and past the bitmap's end its last 16 bytes (rows past the table's
 * end are masked); a bitmap has (rows + 31) / 32 words and 16 bytes of
 * slack after them. */
__device__ __forceinline__ qk_u64 qk_ss_bitmap_src(qk_cu32p prog, qk_cu64p cols, qk_u32 j, qk_u32 lane, qk_u64 base,
                                                   qk_u64 tile, qk_u64 rows_end)
{
  const qk_u32 nb = prog[8];
  const qk_cu32p tab = prog + prog[9];
  const qk_u32 lpb = (qk_u32)(tile >> 7);
  qk_u32 s = j * (64u / lpb) + lane / lpb;
  s = s < nb ? s : nb - 1u;
  const qk_u64 last = ((((rows_end + 31) >> 5) * 4 - 1) & ~15ULL);
  qk_u64 off = (base >> 3) + (qk_u64)(lane % lpb) * 16u;
  off = off <= last ? off : last;
  return cols[tab[2 * s]] + (qk_u64)tab[2 * s + 1] * 16u + off;
}

template <int WG>
__device__ __forceinline__ void qk_ss_stage_issue(qk_cu32p prog, qk_cu32p st, qk_u32 U, qk_cu64p cols, qk_u64 base,
                                                  qk_u64 tile, qk_u64 rows_end, unsigned char* buf)
{
  /* the wave index is uniform; readfirstlane tells the compiler so (the
   * LDS destination of global_load_lds must be uniform) */
  const qk_u32 wave = (qk_u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63u;
  for (qk_u32 u = 0; u < U; u++) {
    const qk_u32 c = st[3 * u], width = st[3 * u + 1], off = st[3 * u + 2];
    const qk_u64 col_bytes = rows_end * width;
    const qk_u64 last = (col_bytes - 1) & ~15ULL;
    const qk_u64 start = base * width;
    const qk_u32 chunks = (qk_u32)((tile * width) >> 10);
    for (qk_u32 k = wave; k < chunks; k += WG / 64) {
      const qk_u64 cb = start + ((qk_u64)k << 10);
      if (cb >= col_bytes) break;
      qk_u64 src = cb + lane * 16u;
      src = src <= last ? src : last;
      __builtin_amdgcn_global_load_lds((const QK_GLOBAL void*)(cols[c] + src),
                                       (__attribute__((address_space(3))) void*)(buf + off + (k << 10)), 16, 0, 2 /* nt */);
    }
  }
  /* the null bitmaps' chunks */
  const qk_u32 nbc = prog[11], area = prog[10];
  for (qk_u32 j = wave; j < nbc; j += WG / 64) {
    const qk_u64 src = qk_ss_bitmap_src(prog, cols, j, lane, base, tile, rows_end);
    __builtin_amdgcn_global_load_lds((const QK_GLOBAL void*)src,
                                     (__attribute__((address_space(3))) void*)(buf + area + (j << 10)), 16, 0, 2 /* nt */);
  }
}

/* the register-staged form (1024-thread blocks): wave w holds chunks w,
 * w + 16, ... of the chunk table in PF registers of 16 bytes a lane.  The
 * loads of tile t + 1 are issued right after tile t is in LDS and land
 * while the waves evaluate tile t; stored into the one stage buffer once
 * every wave is done with tile t.  Lanes past a column's end read its last
 * 16 bytes (those rows are masked). */
template <int WG, int PF>
__device__ __forceinline__ void qk_ss_pf_issue(qk_cu32p prog, qk_cu32p st, qk_cu64p cols, qk_u64 base, qk_u64 rows_end,
                                               qk_u32x4 (&pf)[PF])
{
  const qk_u32 wave = (qk_u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63u;
  const qk_cu32p ct = prog + prog[6];
  const qk_u32 tot = prog[7];
#pragma unroll
  for (int s = 0; s < PF; s++) {
    const qk_u32 f = wave + (qk_u32)s * (WG / 64);
    if (f < tot) {
      const qk_u32 e = ct[f], u = e & 0xffffu, k = e >> 16;
      if (u & 0x8000u) {
        /* chunk k of the null bitmaps (tiles of WG rows) */
        pf[s] = qk_ld(qk_gptr<const qk_u32x4>(qk_ss_bitmap_src(prog, cols, k, lane, base, WG, rows_end)));
        continue;
      }
      const qk_u32 width = st[3 * u + 1];
      const qk_u64 last = ((rows_end * width - 1) & ~15ULL);
      qk_u64 src = base * width + ((qk_u64)k << 10) + lane * 16u;
      src = src <= last ? src : last;
      pf[s] = qk_ld(qk_gptr<const qk_u32x4>(cols[st[3 * u]] + src));
    }
  }
}

template <int WG, int PF>
__device__ __forceinline__ void qk_ss_pf_store(qk_cu32p prog, qk_cu32p st, const qk_u32x4 (&pf)[PF], unsigned char* buf)
{
  const qk_u32 wave = (qk_u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63u;
  const qk_cu32p ct = prog + prog[6];
  const qk_u32 tot = prog[7];
#pragma unroll
  for (int s = 0; s < PF; s++) {
    const qk_u32 f = wave + (qk_u32)s * (WG / 64);
    if (f < tot) {
      const qk_u32 e = ct[f], u = e & 0xffffu, k = e >> 16;
      const qk_u32 dst = (u & 0x8000u) ? prog[10] : st[3 * u + 2];
      *(qk_u32x4*)(buf + dst + (k << 10) + lane * 16u) = pf[s];
    }
  }
}

/* SM: the stage mode fixed at compile time (2, registers) or read from
 * push slot 6 (-1).  An instance without the LDS-direct copies matters:
 * while such copies may be in flight the compiler waits for every
 * outstanding global load before each LDS store or atomic, which in the
 * register-staged form would wait out the next tile's loads at every
 * group-table atomic. */
template <int WG, int ITEMS, int SM = -1>
__device__ __forceinline__ void qk_shared_scan(const QK_Args& a)
{
  /* 16-byte registers a lane holds for the register-staged form (1024-
   * thread blocks: a 160 KB stage is 160 chunks, 10 per wave) */
  constexpr int PF = WG == 1024 ? 10 : 1;
  constexpr int RPL = WG * ITEMS / 64; /* rows per lane of a tile */
  constexpr int HF = RPL > 8 ? 8 : RPL; /* f64 values of a lane's rows live at once */
  static_assert(RPL <= 32, "row masks are 32 bits");
  extern __shared__ __attribute__((aligned(16))) qk_u64 lds_u64[];
  const qk_cu32p prog = (qk_cu32p)a.ptr[0];
  const qk_cu64p cols = (qk_cu64p)a.ptr[2];
  const qk_u32 Q = prog[0], W = prog[1], U = prog[3], stage_bytes = prog[4], WL = prog[5];
  const qk_cu32p ops = prog + prog[2];
  const qk_cu32p st = prog + QK_SS_HDR + Q;
  qk_u64* acc = lds_u64;
  unsigned char* buf0 = (unsigned char*)lds_u64 + (((qk_u64)WL * 8 + 15) & ~15ULL);
  /* the waves take a tile's queries from a counter (in the order the
   * host listed them, most expensive first), so a wave that drew cheap
   * queries takes more instead of waiting at the next tile's barrier;
   * tile t uses counter t & 1, reset for tile t + 1 once every wave has
   * passed tile t's barrier */
  __shared__ qk_u32 s_qnext[2];
  if (threadIdx.x == 0) {
    s_qnext[0] = 0u;
    s_qnext[1] = 0u;
  }
  {
    /* every replica of a query's words starts at its op's identity */
    const qk_u32 wave = (qk_u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63u;
    for (qk_u32 q = wave; q < Q; q += WG / 64) {
      const qk_cu32p rec = prog + prog[QK_SS_HDR + q];
      const qk_u32 A = rec[6] < QK_SS_MAX_ARGS ? rec[6] : QK_SS_MAX_ARGS;
      const qk_u32 Wq = rec[5] * (1u + 3u * A), reps = rec[13];
      const qk_cu32p qops = ops + rec[11];
      for (qk_u32 c = 0; c < reps; c++)
        for (qk_u32 j = lane; j < Wq; j += 64) acc[rec[12] + c * Wq + j] = qops[j] == 2u ? ~0ULL : 0ULL;
    }
  }
  const qk_u64 rows = qk_push(a, QK_P_ROWS);
  const qk_u64 tile = (qk_u64)WG * ITEMS;
  const qk_u64 tiles = (rows + tile - 1) / tile;
  const qk_u64 per = (tiles + gridDim.x - 1) / gridDim.x;
  const qk_u64 t_begin = per * blockIdx.x, t_end = t_begin + per < tiles ? t_begin + per : tiles;
  /* push slot 6: 1 = two stage buffers (tile t+1 copied into LDS while
   * tile t is evaluated), 0 = one (a larger tile in the same LDS; other
   * blocks on the CU cover the load), 2 = one buffer filled from registers
   * (1024-thread blocks whose stage is at most 16 x PF chunks) */
  const qk_u32 smode = SM >= 0 ? (qk_u32)SM : (qk_u32)qk_push(a, QK_P_MODE);
  const bool dbl = smode == 1u;
  const bool regs = WG == 1024 && smode == 2u;
  qk_u32x4 pf[PF];
  if (regs && t_begin < t_end) qk_ss_pf_issue<WG, PF>(prog, st, cols, t_begin * tile, rows, pf);
  if (dbl && t_begin < t_end) qk_ss_stage_issue<WG>(prog, st, U, cols, t_begin * tile, tile, rows, buf0);
  for (qk_u64 t = t_begin; t < t_end; t++) {
    if (regs) {
      /* every wave is done with tile t-1; tile t's registers into LDS */
      if (t != t_begin) __syncthreads();
      qk_ss_pf_store<WG, PF>(prog, st, pf, buf0);
      __syncthreads();
      if (t + 1 < t_end) qk_ss_pf_issue<WG, PF>(prog, st, cols, (t + 1) * tile, rows, pf);
    } else {
      if (!dbl) {
        if (t != t_begin) __syncthreads(); /* every wave is done with tile t-1 */
        qk_ss_stage_issue<WG>(prog, st, U, cols, t * tile, tile, rows, buf0);
      }
      /* tile t has landed in its buffer, and every wave is done with tile
       * t-1, whose buffer the copies of tile t+1 now take */
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
      if (dbl && t + 1 < t_end)
        qk_ss_stage_issue<WG>(prog, st, U, cols, (t + 1) * tile, tile, rows, buf0 + ((t + 1 - t_begin) & 1u) * stage_bytes);
    }
    const unsigned char* sb = buf0 + (dbl ? ((t - t_begin) & 1u) * stage_bytes : 0);
    /* a wave evaluates each query it takes over all rows of the tile:
     * lane l takes rows l, l + 64, ... (RPL rows) */
    const qk_u64 base = t * tile;
    const qk_u32 lane = threadIdx.x & 63u;
    qk_u32 valid = 0;
#pragma unroll
    for (int i = 0; i < RPL; i++) valid |= (base + (qk_u64)i * 64 + lane < rows) ? (1u << i) : 0u;
    const qk_u32 cur = (qk_u32)((t - t_begin) & 1u);
    if (threadIdx.x == 0) s_qnext[cur ^ 1u] = 0u;
    /* the wave takes its next query while it evaluates this one, so the
     * counter's round trip overlaps the work */
    qk_u32 qn = 0u;
    if (lane == 0) qn = atomicAdd(&s_qnext[cur], 1u);
#pragma unroll 1
    for (;;) {
      const qk_u32 q = (qk_u32)__builtin_amdgcn_readfirstlane((int)qn);
      if (q >= Q) break;
      if (lane == 0) qn = atomicAdd(&s_qnext[cur], 1u);
      const qk_cu32p rec = prog + prog[QK_SS_HDR + q];
      const qk_u32 L = (rec[0] & 0xffu) < 8u ? (rec[0] & 0xffu) : 8u;
      const qk_u32 T = (rec[0] >> 8) & 0xffu; /* terms: 0 = the AND of the leaves */
      const qk_u32 flags = rec[1];
      qk_u32 m = valid;
      qk_u32 lmv[8];
#pragma unroll
      for (int k = 0; k < 8; k++) lmv[k] = 0u;
#pragma unroll 1
      for (qk_u32 l = 0; l < L; l++) {
        const qk_cu32p w = rec + QK_SS_REC + 6 * l;
        const qk_u32 u = w[0], tw = w[1], op = w[2], type = tw & 0xffu;
        const unsigned char* cp = sb + u;
        qk_u32 lm = ~0u;
        const bool null_test = (tw & QK_LEAF_NULL_TEST) != 0u; /* IS [NOT] NULL: the bitmap alone */
        if (null_test) {
          lm = op == 5u ? 0u : ~0u; /* a column without NULLs */
        } else if (tw & QK_LEAF_NEVER) {
          lm = 0u;
        } else if ((tw & QK_LEAF_INT) && (type == QK_T_U32 || type == QK_T_I32)) {
          const qk_u32 lo = w[3], bias = (tw & QK_LEAF_BIAS) ? 0x80000000u : 0u;
          const qk_u32* sv = (const qk_u32*)cp;
          if (tw & QK_LEAF_SET) {
            const qk_u64 set = (qk_u64)w[2] | ((qk_u64)w[4] << 32);
#pragma unroll
            for (int i = 0; i < RPL; i++) {
              const qk_u32 d = (sv[i * 64 + lane] ^ bias) - lo;
              if (!(d < 64u && ((set >> d) & 1u))) lm &= ~(1u << i);
            }
          } else {
            const qk_u32 span = w[4];
#pragma unroll
            for (int i = 0; i < RPL; i++)
              if (((sv[i * 64 + lane] ^ bias) - lo) > span) lm &= ~(1u << i);
          }
        } else {
          const double tv = __longlong_as_double((qk_s64)((qk_u64)w[3] | ((qk_u64)w[4] << 32)));
#pragma unroll
          for (int h = 0; h < RPL; h += HF) {
            double x[HF];
            qk_ss_rows<HF>(cp, type, lane, h, x);
            lm = qk_ss_cmp<HF>(x, op, tv, lm, h);
          }
        }
        /* != and NOT IN: the complement (a NULL row still fails below) */
        if (!null_test && (tw & QK_LEAF_NEGATE)) lm = ~lm;
        if (w[5]) {
          /* a nullable column: its NULL rows fail the comparison (IS NULL
           * keeps just them, IS NOT NULL the others).  The tile's part of
           * the bitmap is staged (16 * (w[5] - 1) bytes into the stage):
           * row i * 64 + lane of the tile at bit lane mod 32 of word
           * 2 i + lane / 32; rows past the table's end are outside
           * `valid`. */
          const qk_u32* nb = (const qk_u32*)(sb + (qk_u64)(w[5] - 1u) * 16u);
          qk_u32 nm = 0;
#pragma unroll
          for (int i = 0; i < RPL; i++) nm |= ((nb[2 * i + (lane >> 5)] >> (lane & 31u)) & 1u) << i;
          lm = null_test ? (op == 5u ? nm : ~nm) : (lm & ~nm);
        }
        if (T == 0u) {
          m &= lm;
        } else {
#pragma unroll
          for (int k = 0; k < 8; k++)
            if ((qk_u32)k == l) lmv[k] = lm;
        }
      }
      if (T) {
        /* a disjunction of conjunctions: the terms' leaf bits follow the
         * leaves, a byte each */
        const qk_cu32p tw4 = rec + QK_SS_REC + 6 * L;
        qk_u32 r = 0u;
#pragma unroll 1
        for (qk_u32 t = 0; t < T && t < 16u; t++) {
          const qk_u32 tm = (tw4[t >> 2] >> (8u * (t & 3u))) & 0xffu;
          qk_u32 c = valid;
#pragma unroll
          for (int k = 0; k < 8; k++)
            if (tm & (1u << k)) c &= lmv[k];
          r |= c;
        }
        m = r;
      }
      const qk_u32 R = rec[5], A = rec[6] < QK_SS_MAX_ARGS ? rec[6] : QK_SS_MAX_ARGS;
      qk_u64* qa = acc + rec[12];
      if (flags & 1u) {
        /* grouped: LDS atomics into the query's group table (only this
         * wave writes it); lane l adds into replica l mod reps, so lanes
         * whose rows share a group contend reps times less.  The rows are
         * walked in a fixed unrolled order under a per-row predicate, with
         * the group of every row and the argument values read first. */
        if (!__any((int)(m != 0))) continue;
        qa += (qk_u64)(lane & (rec[13] - 1u)) * (R * (1u + 3u * A));
        const qk_u32* kp = (const qk_u32*)(sb + rec[2]);
        const qk_u32 kx = rec[3] == QK_T_I32 ? 0x80000000u : 0u, kmin = rec[4];
        if ((flags & 8u) && !__any((int)(__popc(m) > (qk_u32)RPL / 4u))) {
          /* compact, for a wave whose lanes each hold at most RPL / 4
           * passing rows: each lane walks only those, lowest first, so the
           * wave loops as often as its busiest lane has rows rather than
           * RPL times per key, count and argument (with more rows the
           * unrolled walk below, which overlaps its LDS reads, is faster) */
          qk_u32 mm = m;
          while (mm) {
            const qk_u32 r = (qk_u32)__builtin_ctz(mm) * 64u + lane;
            mm &= mm - 1u;
            const qk_u32 gi = (kp[r] ^ kx) - kmin;
            if (gi >= R) continue;
            atomicAdd((unsigned long long*)&qa[gi], 1ULL);
#pragma unroll 1
            for (qk_u32 k = 0; k < A; k++) {
              const qk_u32 at = rec[8 + 2 * k];
              const unsigned char* ap = sb + rec[7 + 2 * k];
              qk_u64* qs = qa + R + (qk_u64)k * 3 * R;
              qk_u64 add;
              double v;
              if (qk_type_is_int32(at)) {
                const qk_u32 x = ((const qk_u32*)ap)[r];
                const bool sgn = at == QK_T_I32;
                add = sgn ? (qk_u64)(qk_s64)(qk_s32)x : (qk_u64)x;
                v = sgn ? (double)(qk_s32)x : (double)x;
                if (flags & 2u) atomicAdd((unsigned long long*)&qs[gi], add);
              } else {
                v = qk_ss_f64(ap, at, r);
                if (flags & 2u) atomicAdd((double*)&qs[gi], v);
              }
              if (flags & 4u) {
                const qk_u64 kk = qk_f64_key(v);
                atomicMin((unsigned long long*)&qs[R + gi], (unsigned long long)kk);
                atomicMax((unsigned long long*)&qs[2 * R + gi], (unsigned long long)kk);
              }
            }
          }
          continue;
        }
        /* the lane's rows in groups of HF: each group's keys, then its
         * count and argument atomics, so only HF group indices are live */
#pragma unroll
        for (int h = 0; h < RPL; h += HF) {
          qk_u32 g[HF];
#pragma unroll
          for (int i = 0; i < HF; i++) {
            g[i] = (kp[(h + i) * 64 + lane] ^ kx) - kmin;
            if (g[i] >= R) m &= ~(1u << (h + i));
          }
#pragma unroll
          for (int i = 0; i < HF; i++)
            if ((m >> (h + i)) & 1u) atomicAdd((unsigned long long*)&qa[g[i]], 1ULL);
#pragma unroll 1
          for (qk_u32 k = 0; k < A; k++) {
            const qk_u32 at = rec[8 + 2 * k];
            const unsigned char* ap = sb + rec[7 + 2 * k];
            qk_u64* qs = qa + R + (qk_u64)k * 3 * R;
            if (qk_type_is_int32(at)) {
              /* 32-bit integers: exact sums without a trip through f64 */
              const qk_u32* av = (const qk_u32*)ap;
              const bool sgn = at == QK_T_I32;
              if (flags & 2u) {
#pragma unroll
                for (int i = 0; i < HF; i++) {
                  const qk_u32 x = av[(h + i) * 64 + lane];
                  if ((m >> (h + i)) & 1u)
                    atomicAdd((unsigned long long*)&qs[g[i]], sgn ? (qk_u64)(qk_s64)(qk_s32)x : (qk_u64)x);
                }
              }
              if (flags & 4u) {
#pragma unroll
                for (int i = 0; i < HF; i++) {
                  const qk_u32 x = av[(h + i) * 64 + lane];
                  const qk_u64 kk = qk_f64_key(sgn ? (double)(qk_s32)x : (double)x);
                  if ((m >> (h + i)) & 1u) {
                    atomicMin((unsigned long long*)&qs[R + g[i]], (unsigned long long)kk);
                    atomicMax((unsigned long long*)&qs[2 * R + g[i]], (unsigned long long)kk);
                  }
                }
              }
            } else {
              double v[HF];
              qk_ss_rows<HF>(ap, at, lane, h, v);
              if (flags & 2u) {
#pragma unroll
                for (int i = 0; i < HF; i++)
                  if ((m >> (h + i)) & 1u) atomicAdd((double*)&qs[g[i]], v[i]);
              }
              if (flags & 4u) {
#pragma unroll
                for (int i = 0; i < HF; i++) {
                  const qk_u64 kk = qk_f64_key(v[i]);
                  if ((m >> (h + i)) & 1u) {
                    atomicMin((unsigned long long*)&qs[R + g[i]], (unsigned long long)kk);
                    atomicMax((unsigned long long*)&qs[2 * R + g[i]], (unsigned long long)kk);
                  }
                }
              }
            }
          }
        }
      } else {
        /* ungrouped: the lane's rows in registers, then the wave (DPP),
         * then lane 0 into the query's words (only this wave writes them) */
        if (!__any((int)(m != 0))) continue;
        const qk_u32 cnt = qk_wave_add_u32((qk_u32)__popc(m)); /* at most 64 x 32 */
        if (lane == 0) qa[0] += cnt;
#pragma unroll 1
        for (qk_u32 k = 0; k < A; k++) {
          const qk_u32 at = rec[8 + 2 * k];
          const unsigned char* ap = sb + rec[7 + 2 * k];
          qk_u64* qs = qa + 1 + (qk_u64)k * 3;
          qk_u64 mn = ~0ULL, mx = 0ULL;
          if (qk_type_is_int32(at)) {
            /* 32-bit integers: exact integer sums, and the lane's min/max
             * compared as integers (the f64 key is monotonic in the value)
             * and keyed once */
            const qk_u32* av = (const qk_u32*)ap;
            const bool sgn = at == QK_T_I32;
            const qk_u32 flip = sgn ? 0x80000000u : 0u;
            if (flags & 2u) {
              qk_u64 su = 0;
#pragma unroll
              for (int i = 0; i < RPL; i++) {
                const qk_u32 x = av[i * 64 + lane];
                su += ((m >> i) & 1u) ? (sgn ? (qk_u64)(qk_s64)(qk_s32)x : (qk_u64)x) : 0ULL;
              }
              su = qk_wave_add_u64(su);
              if (lane == 0) qs[0] += su;
            }
            if (flags & 4u) {
              /* order codes (value ^ flip) reduced as 32-bit integers; some
               * lane of the wave has a row here, so the wave's codes are
               * real values, keyed once */
              qk_u32 lo = ~0u, hi = 0u;
#pragma unroll
              for (int i = 0; i < RPL; i++) {
                const qk_u32 c = av[i * 64 + lane] ^ flip;
                const bool in = (m >> i) & 1u;
                lo = (in && c < lo) ? c : lo;
                hi = (in && c > hi) ? c : hi;
              }
              qk_wave_minmax_u32(lo, hi);
              if (lane == 0) {
                const qk_u64 kn = qk_f64_key(sgn ? (double)(qk_s32)(lo ^ flip) : (double)lo);
                const qk_u64 kx = qk_f64_key(sgn ? (double)(qk_s32)(hi ^ flip) : (double)hi);
                qs[1] = kn < qs[1] ? kn : qs[1];
                qs[2] = kx > qs[2] ? kx : qs[2];
              }
              continue;
            }
          } else {
            /* the lane's rows in order, at most HF values live at once */
            double sd = 0.0;
#pragma unroll
            for (int h = 0; h < RPL; h += HF) {
              double v[HF];
              qk_ss_rows<HF>(ap, at, lane, h, v);
              if (flags & 2u) {
#pragma unroll
                for (int i = 0; i < HF; i++) sd += (m & (1u << (h + i))) ? v[i] : 0.0;
              }
              if (flags & 4u) {
#pragma unroll
                for (int i = 0; i < HF; i++) {
                  const qk_u64 kk = qk_f64_key(v[i]);
                  const bool in = (m >> (h + i)) & 1u;
                  mn = (in && kk < mn) ? kk : mn;
                  mx = (in && kk > mx) ? kk : mx;
                }
              }
            }
            if (flags & 2u) {
              sd = qk_wave_add_f64(sd);
              if (lane == 0) qs[0] = (qk_u64)__double_as_longlong(__longlong_as_double((qk_s64)qs[0]) + sd);
            }
          }
          if (flags & 4u) {
            qk_wave_minmax_u64(mn, mx);
            if (lane == 0) {
              qs[1] = mn < qs[1] ? mn : qs[1];
              qs[2] = mx > qs[2] ? mx : qs[2];
            }
          }
        }
      }
    }
  }
  __syncthreads();
  /* the block's words: each query's replicas folded with the word's op */
  qk_u64* out = qk_wptr<qk_u64>(a, 1) + (qk_u64)blockIdx.x * W;
  const qk_u32 fwave = (qk_u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), flane = threadIdx.x & 63u;
  for (qk_u32 q = fwave; q < Q; q += WG / 64) {
    const qk_cu32p rec = prog + prog[QK_SS_HDR + q];
    const qk_u32 A = rec[6] < QK_SS_MAX_ARGS ? rec[6] : QK_SS_MAX_ARGS;
    const qk_u32 Wq = rec[5] * (1u + 3u * A), reps = rec[13], o = rec[11];
    const qk_u64* qa = acc + rec[12];
    for (qk_u32 j = flane; j < Wq; j += 64) {
      const qk_u32 op = ops[o + j];
      qk_u64 x = qa[j];
      for (qk_u32 c = 1; c < reps; c++) {
        const qk_u64 y = qa[c * Wq + j];
        if (op == 0u) x += y;
        else if (op == 1u) x = (qk_u64)__double_as_longlong(__longlong_as_double((qk_s64)x) + __longlong_as_double((qk_s64)y));
        else if (op == 2u) x = y < x ? y : x;
        else x = y > x ? y : x;
      }
      out[o + j] = x;
    }
  }
}

/* instances: shared_scan_<block size>_<rows per thread staged> */
extern "C" __global__ void LB k_shared_scan_256_1(QK_Args a) { qk_shared_scan<256, 1>(a); }
extern "C" __global__ void LB k_shared_scan_256_2(QK_Args a) { qk_shared_scan<256, 2>(a); }
extern "C" __global__ void LB k_shared_scan_256_4(QK_Args a) { qk_shared_scan<256, 4>(a); }
extern "C" __global__ void __launch_bounds__(512) k_shared_scan_512_1(QK_Args a) { qk_shared_scan<512, 1>(a); }
extern "C" __global__ void __launch_bounds__(512) k_shared_scan_512_2(QK_Args a) { qk_shared_scan<512, 2>(a); }
extern "C" __global__ void __launch_bounds__(1024) k_shared_scan_1024_1(QK_Args a) { qk_shared_scan<1024, 1>(a); }
/* the register-staged form alone (stage mode 2) */
extern "C" __global__ void __launch_bounds__(1024) k_shared_scan_1024_1_r(QK_Args a) { qk_shared_scan<1024, 1, 2>(a); }

/* shared_reduce: word w of the result folds the blocks' words w with the
 * word's op.  One wave per word: lane l folds blocks l, l + 64, ... in
 * order, then the wave combines its lanes (DPP tree), so an f64 sum has
 * one fixed order for a given grid.  binding 0 program, 1 partials
 * [blocks][W], 3 result u64[W]; push 0 blocks */
extern "C" __global__ void LB k_shared_reduce(QK_Args a)
{
  const qk_cu32p prog = (qk_cu32p)a.ptr[0];
  const qk_u32 W = prog[1];
  const qk_cu32p ops = prog + prog[2];
  const qk_u64 nb = qk_push(a, 0);
  const qk_u64* part = qk_ptr<qk_u64>(a, 1);
  qk_u64* res = qk_wptr<qk_u64>(a, 3);
  const qk_u32 wave = (qk_u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63u;
  const qk_u32 w = blockIdx.x * (QK_WG / 64) + wave;
  if (w >= W) return;
  const qk_u32 op = ops[w];
  qk_u64 x = op == 2u ? ~0ULL : 0ULL;
  double d = 0.0;
  for (qk_u64 b = lane; b < nb; b += 64) {
    const qk_u64 v = part[b * W + w];
    if (op == 0u) x += v;
    else if (op == 1u) d += __longlong_as_double((qk_s64)v);
    else if (op == 2u) x = v < x ? v : x;
    else x = v > x ? v : x;
  }
  qk_u64 out;
  if (op == 0u) {
    out = qk_wave_add_u64(x);
  } else if (op == 1u) {
    out = (qk_u64)__double_as_longlong(qk_wave_add_f64(d));
  } else {
    qk_u64 mn = x, mx = x;
    qk_wave_minmax_u64(mn, mx);
    out = op == 2u ? mn : mx;
  }
  if (lane == 0) res[w] = out;
}

/* ------------------------------------------------------------------ */
/* column encoding (qe_encode_chunk / qe_column_is_sequence,            */
/* query_exec.c:1506-1672), on the device from the resident column.     */
/* encode_column: binding 6 the u32/i32 column, 2 codes out (u8 or u16, */
/* 4 rows a quad), 3 flags (u32: set when a value - base exceeds the    */
/* width's range); push 0 rows; words ARGS = width (1, 2), KEY_MIN =    */
/* base.  seq_check: binding 6 the column, 3 flags (set when a row's    */
/* value is not base + row); push 0 rows; word KEY_MIN = base.          */
/* ------------------------------------------------------------------ */
extern "C" __global__ void LB k_encode_column(QK_Args a)
{
  const qk_u64 rows = qk_push(a, 0);
  const qk_u32 width = a.words[QK_W_ARGS], base = a.words[QK_W_KEY_MIN];
  const qk_u32 maxc = width == 1 ? 0xffu : 0xffffu;
  const qk_u32x4* col = qk_ptr<qk_u32x4>(a, QK_B_COL0);
  qk_u32* flags = qk_wptr<qk_u32>(a, 3);
  const qk_u64 quads = (rows + 3) / 4;
  bool bad = false;
  for (qk_u64 q = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; q < quads; q += (qk_u64)gridDim.x * QK_WG) {
    const qk_u32x4 v = col[q];
    const qk_u32 w[4] = {v.x, v.y, v.z, v.w};
    qk_u32 c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const bool live = q * 4 + k < rows;
      c[k] = live ? w[k] - base : 0u;
      bad |= live && c[k] > maxc;
    }
    if (width == 1) qk_wptr<qk_u32>(a, 2)[q] = c[0] | (c[1] << 8) | (c[2] << 16) | (c[3] << 24);
    else qk_wptr<uint2>(a, 2)[q] = make_uint2(c[0] | (c[1] << 16), c[2] | (c[3] << 16));
  }
  if (bad) atomicOr(flags, 1u);
}

extern "C" __global__ void LB k_seq_check(QK_Args a)
{
  const qk_u64 rows = qk_push(a, 0);
  const qk_u32 base = a.words[QK_W_KEY_MIN];
  const qk_u32x4* col = qk_ptr<qk_u32x4>(a, QK_B_COL0);
  const qk_u64 quads = (rows + 3) / 4;
  bool bad = false;
  for (qk_u64 q = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; q < quads; q += (qk_u64)gridDim.x * QK_WG) {
    const qk_u32x4 v = col[q];
    const qk_u32 w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; k++) bad |= q * 4 + k < rows && w[k] != base + (qk_u32)(q * 4 + k);
  }
  if (bad) atomicOr(qk_wptr<qk_u32>(a, 3), 1u);
}

/* ------------------------------------------------------------------ */
/* gather: column values through row ids (selection_gather.comp modes   */
/* F32/F64): out_c[i] = column c at rows[i * STRIDE + OFFSET], 0 where  */
/* the id is all ones (no row).  binding 0 row ids (u64), 6..9 columns, */
/* 10..13 outputs (the column's width); push 0 count, 1 stride, 2       */
/* offset (u64 units); words ARGS = columns (1..4), TYPES = widths,     */
/* 8 bits each (4 or 8 bytes).                                          */
/* ------------------------------------------------------------------ */
extern "C" __global__ void LB k_gather(QK_Args a)
{
  const qk_u64 n = qk_push(a, 0), stride = qk_push(a, 1), offset = qk_push(a, 2);
  const qk_u32 ncols = a.words[QK_W_ARGS] & 0xffu;
  const qk_u32 widths = a.words[QK_W_TYPES];
  const qk_u64* rows = qk_ptr<qk_u64>(a, 0);
  for (qk_u64 i = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; i < n; i += (qk_u64)gridDim.x * QK_WG) {
    const qk_u64 id = rows[i * stride + offset];
#pragma unroll
    for (qk_u32 c = 0; c < 4; c++) {
      if (c >= ncols) break;
      if (((widths >> (8 * c)) & 0xffu) == 8) {
        qk_wptr<qk_u64>(a, 10 + c)[i] = id == ~0ULL ? 0ULL : qk_ptr<qk_u64>(a, QK_B_COL0 + c)[id];
      } else {
        qk_wptr<qk_u32>(a, 10 + c)[i] = id == ~0ULL ? 0u : qk_ptr<qk_u32>(a, QK_B_COL0 + c)[id];
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* exclusive scan of u32 counts (any length): scan_u32_blocks scans each */
/* 2048-element tile and leaves its total, scan_u32_top scans the tile  */
/* totals in one block, scan_u32_add adds them back.  binding 0 in, 1   */
/* out (may alias in), 2 tile totals; push 0 = n.                       */
/* ------------------------------------------------------------------ */
#define QK_SCAN_TILE (QK_WG * 8)
extern "C" __global__ void LB k_scan_u32_blocks(QK_Args a)
{
  __shared__ qk_u32 s_scan[QK_WG / 64];
  const qk_u64 n = qk_push(a, 0);
  const qk_u32* in = qk_ptr<qk_u32>(a, 0);
  qk_u32* out = qk_wptr<qk_u32>(a, 1);
  const qk_u64 t0 = (qk_u64)blockIdx.x * QK_SCAN_TILE + (qk_u64)threadIdx.x * 8;
  qk_u32 v[8], run = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    v[k] = t0 + k < n ? in[t0 + k] : 0u;
    const qk_u32 x = v[k];
    v[k] = run;
    run += x;
  }
  qk_u32 total = 0;
  const qk_u32 pre = qk_block_excl_scan_u32(run, s_scan, total);
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (t0 + k < n) out[t0 + k] = pre + v[k];
  if (threadIdx.x == 0) qk_wptr<qk_u32>(a, 2)[blockIdx.x] = total;
}

extern "C" __global__ void LB k_scan_u32_top(QK_Args a)
{
  __shared__ qk_u32 s_scan[QK_WG / 64];
  const qk_u64 tiles = qk_push(a, 0);
  qk_u32* t = qk_wptr<qk_u32>(a, 2);
  qk_u32 running = 0;
  for (qk_u64 c = 0; c < tiles; c += QK_WG) {
    const qk_u64 i = c + threadIdx.x;
    const qk_u32 v = i < tiles ? t[i] : 0u;
    qk_u32 total = 0;
    const qk_u32 pre = qk_block_excl_scan_u32(v, s_scan, total);
    if (i < tiles) t[i] = running + pre;
    running += total;
  }
}

extern "C" __global__ void LB k_scan_u32_add(QK_Args a)
{
  const qk_u64 n = qk_push(a, 0);
  qk_u32* out = qk_wptr<qk_u32>(a, 1);
  const qk_u32 add = qk_ptr<qk_u32>(a, 2)[blockIdx.x];
  const qk_u64 t0 = (qk_u64)blockIdx.x * QK_SCAN_TILE;
  for (qk_u32 k = threadIdx.x; k < QK_SCAN_TILE; k += QK_WG)
    if (t0 + k < n) out[t0 + k] += add;
}

/* ------------------------------------------------------------------ */
/* hash join producing row pairs (query_exec.c:14541-15358,             */
/* hash_join_build_count.comp / csr_scatter.comp / hash_join_probe.comp) */
/* Keys compare as f64 values (u32/i32/u64/i64/f32/f64 columns), -0.0   */
/* as 0.0, NaN never equal.  The build side goes into an open-addressing */
/* table of distinct key codes (2^LOG2 >= 2 x build rows, linear        */
/* probing from fmix64(code)) with a count per key; a scan of the counts */
/* and a scatter give each key's build positions as one contiguous      */
/* group, which is then sorted, so a probe row's matches come out in    */
/* build-position order.                                                */
/* Common bindings: 2 table keys u64[T], 3 counts u32[T], 5 group starts */
/* u32[T], 8 groups u32[build rows]; word LOG2.                         */
/* ------------------------------------------------------------------ */
#define QK_HJ_NONE 0xffffffffu

__device__ __forceinline__ qk_u32 qk_hj_lookup(const qk_u64* keys, qk_u64 code, qk_u64 mask)
{
  qk_u64 s = qk_fmix64(code) & mask;
  for (qk_u64 i = 0; i <= mask; i++) {
    const qk_u64 k = keys[s];
    if (k == code) return (qk_u32)s;
    if (k == QK_HASH_EMPTY) return QK_HJ_NONE;
    s = (s + 1) & mask;
  }
  return QK_HJ_NONE;
}

/* hjoin_build: binding 0 build position -> row (u64, optional), 4 slot of
 * each build position (u32, NONE for a NaN key), 6 build key column; push
 * 0 = build positions; word KEY = key column type.  Keys start all-ones
 * (never a code: it is a NaN pattern), counts at zero. */
extern "C" __global__ void LB k_hjoin_build(QK_Args a)
{
  const qk_u64 nb = qk_push(a, 0);
  const qk_u32 type = a.words[QK_W_KEY] & 0xffu;
  const qk_u64 mask = (1ULL << a.words[QK_W_LOG2]) - 1;
  const qk_u64* list = a.ptr[0] ? qk_ptr<qk_u64>(a, 0) : nullptr;
  qk_u64* keys = qk_wptr<qk_u64>(a, 2);
  qk_u32* cnt = qk_wptr<qk_u32>(a, 3);
  qk_u32* slot_of = qk_wptr<qk_u32>(a, 4);
  for (qk_u64 i = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; i < nb; i += (qk_u64)gridDim.x * QK_WG) {
    const double d = qk_load1_f64(a, 0, type, list ? list[i] : i);
    qk_u32 slot = QK_HJ_NONE;
    if (d == d) {
      const qk_u64 code = qk_key_code_f64(d);
      qk_u64 s = qk_fmix64(code) & mask;
      for (qk_u64 t = 0; t <= mask; t++) {
        const qk_u64 prev = atomicCAS((unsigned long long*)&keys[s], (unsigned long long)QK_HASH_EMPTY,
                                      (unsigned long long)code);
        if (prev == QK_HASH_EMPTY || prev == code) {
          atomicAdd(&cnt[s], 1u);
          slot = (qk_u32)s;
          break;
        }
        s = (s + 1) & mask;
      }
    }
    slot_of[i] = slot;
  }
}

/* hjoin_scatter: build position i goes to groups[start[slot] + rank], the
 * rank from an atomic cursor per slot (binding 7, zeroed). */
extern "C" __global__ void LB k_hjoin_scatter(QK_Args a)
{
  const qk_u64 nb = qk_push(a, 0);
  const qk_u32* slot_of = qk_ptr<qk_u32>(a, 4);
  const qk_u32* start = qk_ptr<qk_u32>(a, 5);
  qk_u32* cursor = qk_wptr<qk_u32>(a, 7);
  qk_u32* groups = qk_wptr<qk_u32>(a, 8);
  for (qk_u64 i = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; i < nb; i += (qk_u64)gridDim.x * QK_WG) {
    const qk_u32 s = slot_of[i];
    if (s == QK_HJ_NONE) continue;
    groups[start[s] + atomicAdd(&cursor[s], 1u)] = (qk_u32)i;
  }
}

/* hjoin_sort_small: one thread per slot sorts a group of 2..32 positions
 * in place (insertion sort); larger groups are listed in binding 9 (u32,
 * count in binding 10 [0], zeroed) for hjoin_sort_big.  push 0 = T. */
#define QK_HJ_SMALL 32
extern "C" __global__ void LB k_hjoin_sort_small(QK_Args a)
{
  const qk_u64 T = qk_push(a, 0);
  const qk_u32* cnt = qk_ptr<qk_u32>(a, 3);
  const qk_u32* start = qk_ptr<qk_u32>(a, 5);
  qk_u32* groups = qk_wptr<qk_u32>(a, 8);
  qk_u32* big = qk_wptr<qk_u32>(a, 9);
  qk_u32* big_count = qk_wptr<qk_u32>(a, 10);
  for (qk_u64 s = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; s < T; s += (qk_u64)gridDim.x * QK_WG) {
    const qk_u32 c = cnt[s];
    if (c < 2) continue;
    if (c > QK_HJ_SMALL) {
      big[atomicAdd(big_count, 1u)] = (qk_u32)s;
      continue;
    }
    qk_u32* g = groups + start[s];
    qk_u32 v[QK_HJ_SMALL];
    for (qk_u32 k = 0; k < c; k++) v[k] = g[k];
    for (qk_u32 k = 1; k < c; k++) {
      const qk_u32 x = v[k];
      qk_u32 j = k;
      while (j > 0 && v[j - 1] > x) {
        v[j] = v[j - 1];
        j--;
      }
      v[j] = x;
    }
    for (qk_u32 k = 0; k < c; k++) g[k] = v[k];
  }
}

/* hjoin_sort_big: block b sorts the groups listed at b, b + grid, ... with
 * a bitonic network in which every compare-exchange puts the smaller value
 * at the lower index (each merge starts by comparing mirrored positions),
 * so positions past the group's end act as +infinity and are never read
 * or written. */
extern "C" __global__ void LB k_hjoin_sort_big(QK_Args a)
{
  const qk_u32* cnt = qk_ptr<qk_u32>(a, 3);
  const qk_u32* start = qk_ptr<qk_u32>(a, 5);
  qk_u32* groups = qk_wptr<qk_u32>(a, 8);
  const qk_u32* big = qk_ptr<qk_u32>(a, 9);
  const qk_u32 nbig = qk_ptr<qk_u32>(a, 10)[0];
  for (qk_u32 gi = blockIdx.x; gi < nbig; gi += gridDim.x) {
    const qk_u32 s = big[gi];
    const qk_u32 c = cnt[s];
    qk_u32* g = groups + start[s];
    qk_u32 N = 1;
    while (N < c) N <<= 1;
    for (qk_u32 half = 1; half < N; half <<= 1) {
      for (qk_u32 dist = half; dist > 0; dist >>= 1) {
        __syncthreads();
        for (qk_u32 i = threadIdx.x; i < N; i += QK_WG) {
          /* first stage of a merge: mirror within the 2*half block */
          const qk_u32 l = dist == half ? (i ^ (2 * half - 1)) : (i ^ dist);
          if (l <= i || l >= c) continue;
          const qk_u32 x = g[i], y = g[l];
          if (x > y) {
            g[i] = y;
            g[l] = x;
          }
        }
      }
    }
    __syncthreads();
  }
}

/* hjoin_probe: binding 0 probe position -> row (u64, optional), 1 build
 * position -> row (u64, optional), 6 probe key column, 9 out records
 * (u64), 10 per-block totals (mode 0) / first record of each block (mode
 * 1); push 0 = probe positions, 1 = mode, 2 = record capacity; words KEY
 * = probe key type, ARGS = join type (0 inner, 1 left, 2 semi, 3 anti).
 * Each block owns a contiguous range of probe positions.  inner/left
 * records are (probe position, build row) pairs, a left join's unmatched
 * row paired with all ones; semi/anti records are probe positions. */
__device__ __forceinline__ qk_u32 qk_hj_outputs(qk_u32 jt, qk_u32 c)
{
  return jt == 0 ? c : (jt == 1 ? (c ? c : 1u) : (jt == 2 ? (c ? 1u : 0u) : (c ? 0u : 1u)));
}

extern "C" __global__ void LB k_hjoin_probe(QK_Args a)
{
  __shared__ qk_u64 s_red[QK_WG / 64];
  __shared__ qk_u32 s_scan[QK_WG / 64];
  const qk_u64 np = qk_push(a, 0);
  const qk_u32 mode = (qk_u32)qk_push(a, 1);
  const qk_u64 cap = qk_push(a, 2);
  const qk_u32 type = a.words[QK_W_KEY] & 0xffu;
  const qk_u32 jt = a.words[QK_W_ARGS] & 0xffu;
  const qk_u64 mask = (1ULL << a.words[QK_W_LOG2]) - 1;
  const qk_u64* plist = a.ptr[0] ? qk_ptr<qk_u64>(a, 0) : nullptr;
  const qk_u64* blist = a.ptr[1] ? qk_ptr<qk_u64>(a, 1) : nullptr;
  const qk_u64* keys = qk_ptr<qk_u64>(a, 2);
  const qk_u32* cnt = qk_ptr<qk_u32>(a, 3);
  const qk_u32* start = qk_ptr<qk_u32>(a, 5);
  const qk_u32* groups = qk_ptr<qk_u32>(a, 8);
  qk_u64* out = qk_wptr<qk_u64>(a, 9);
  qk_u64* blk = qk_wptr<qk_u64>(a, 10);
  qk_u64 per = (np + gridDim.x - 1) / gridDim.x;
  per = (per + QK_WG - 1) / QK_WG * QK_WG;
  const qk_u64 p0 = min(np, per * blockIdx.x), p1 = min(np, p0 + per);
  qk_u64 total = 0, o = mode ? blk[blockIdx.x] : 0;
  for (qk_u64 base = p0; base < p1; base += QK_WG) {
    const qk_u64 p = base + threadIdx.x;
    qk_u32 slot = QK_HJ_NONE, c = 0;
    if (p < p1) {
      const double d = qk_load1_f64(a, 0, type, plist ? plist[p] : p);
      if (d == d) slot = qk_hj_lookup(keys, qk_key_code_f64(d), mask);
      c = slot != QK_HJ_NONE ? cnt[slot] : 0u;
    }
    const qk_u32 n_out = p < p1 ? qk_hj_outputs(jt, c) : 0u;
    if (!mode) {
      total += n_out;
      continue;
    }
    qk_u32 step_total = 0;
    const qk_u32 pre = qk_block_excl_scan_u32(n_out, s_scan, step_total);
    const qk_u64 w = o + pre;
    if (jt >= 2) {
      if (n_out && w < cap) out[w] = p;
    } else if (c) {
      const qk_u32 g0 = start[slot];
      for (qk_u32 k = 0; k < c && w + k < cap; k++) {
        const qk_u32 bpos = groups[g0 + k];
        out[2 * (w + k)] = p;
        out[2 * (w + k) + 1] = blist ? blist[bpos] : (qk_u64)bpos;
      }
    } else if (n_out && w < cap) {
      out[2 * w] = p;
      out[2 * w + 1] = ~0ULL;
    }
    o += step_total;
  }
  if (!mode) {
    const qk_u64 t = qk_block_sum_u64(total, s_red);
    if (threadIdx.x == 0) blk[blockIdx.x] = t;
  }
}

/* ------------------------------------------------------------------ */
/* generators: splitmix64 addressed by output index                     */
/* (bench_common.h:30-37): d(k) = mix(seed + (k+1) * 0x9E3779B97F4A7C15) */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ qk_u64 qk_splitmix64_at(qk_u64 seed, qk_u64 k)
{
  qk_u64 z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* gen_events: bindings 6 id, 7 region, 8 amount, 9 ts (absent = skip);
 * push 0 rows, 1 seed, 2 first row.  region = d(3r)%50,
 * amount = d(3r+1)%10000, ts = d(3r+2)>>32, id = r */
extern "C" __global__ void LB k_gen_events(QK_Args a)
{
  qk_u64 n = qk_push(a, 0), seed = qk_push(a, 1), row0 = qk_push(a, 2);
  qk_u32* id = qk_gptr<qk_u32>(a.ptr[6]);
  qk_u32* region = qk_gptr<qk_u32>(a.ptr[7]);
  qk_u32* amount = qk_gptr<qk_u32>(a.ptr[8]);
  qk_u32* ts = qk_gptr<qk_u32>(a.ptr[9]);
  for (qk_u64 i = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; i < n; i += (qk_u64)gridDim.x * QK_WG) {
    qk_u64 r = row0 + i;
    if (id) id[i] = (qk_u32)r;
    if (region) region[i] = (qk_u32)(qk_splitmix64_at(seed, 3 * r) % 50ULL);
    if (amount) amount[i] = (qk_u32)(qk_splitmix64_at(seed, 3 * r + 1) % 10000ULL);
    if (ts) ts[i] = (qk_u32)(qk_splitmix64_at(seed, 3 * r + 2) >> 32);
  }
}

/* gen_scale_events: the closed-form table of bench_select_scale.c:39-45;
 * bindings 6 id, 7 region, 8 amount, 9 needle key; push 0 rows, 2 first row */
extern "C" __global__ void LB k_gen_scale_events(QK_Args a)
{
  qk_u64 n = qk_push(a, 0), row0 = qk_push(a, 2);
  qk_u32* id = qk_gptr<qk_u32>(a.ptr[6]);
  qk_u32* region = qk_gptr<qk_u32>(a.ptr[7]);
  qk_u32* amount = qk_gptr<qk_u32>(a.ptr[8]);
  qk_u32* key = qk_gptr<qk_u32>(a.ptr[9]);
  for (qk_u64 i = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; i < n; i += (qk_u64)gridDim.x * QK_WG) {
    qk_u64 r = row0 + i;
    if (id) id[i] = (qk_u32)(r + 1);
    if (region) region[i] = (qk_u32)(r % 50);
    if (amount) amount[i] = (qk_u32)(r % 10000);
    if (key) key[i] = (qk_u32)(r * 2654435761ULL);
  }
}

/* gen_fact_fk: fk[r] = dim_keys[d'(r) % dim_rows], d' seeded with
 * seed ^ 0x5851F42D4C957F2D; binding 6 fk out, 0 dim keys;
 * push 0 rows, 1 seed, 2 first row, 3 dim rows */
extern "C" __global__ void LB k_gen_fact_fk(QK_Args a)
{
  qk_u64 n = qk_push(a, 0), seed = qk_push(a, 1) ^ 0x5851F42D4C957F2DULL, row0 = qk_push(a, 2), dn = qk_push(a, 3);
  qk_u32* fk = qk_gptr<qk_u32>(a.ptr[6]);
  const qk_u32* dk = qk_ptr<qk_u32>(a, 0);
  for (qk_u64 i = (qk_u64)blockIdx.x * QK_WG + threadIdx.x; i < n; i += (qk_u64)gridDim.x * QK_WG) {
    fk[i] = dk[qk_splitmix64_at(seed, row0 + i) % dn];
  }
}

/* zonemap_build: per zone of 2^shift rows, min and max as f64
 * (CQL_ZoneMapChunk, cql.h:156-162).  binding 6 column, 1 mins, 2 maxs;
 * push 0 rows, words[2] type, words[3] shift.  One block per zone. */
extern "C" __global__ void LB k_zonemap_build(QK_Args a)
{
  __shared__ double s_mn[QK_WG / 64], s_mx[QK_WG / 64];
  qk_u64 n = qk_push(a, 0);
  qk_u32 type = a.words[QK_W_ZTYPE], shift = a.words[QK_W_ZSHIFT];
  qk_u64 zones = (n + (1ULL << shift) - 1) >> shift;
  for (qk_u64 z = blockIdx.x; z < zones; z += gridDim.x) {
    qk_u64 z0 = z << shift, z1 = min(n, (z + 1) << shift);
    double mn = 1.7976931348623157e308, mx = -1.7976931348623157e308;
    for (qk_u64 r = z0 + threadIdx.x; r < z1; r += QK_WG) {
      double v = qk_load1_f64(a, 0, type, r);
      mn = min(mn, v);
      mx = max(mx, v);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mn = min(mn, __shfl_xor(mn, o, 64));
      mx = max(mx, __shfl_xor(mx, o, 64));
    }
    __syncthreads();
    if (__lane_id() == 0) {
      s_mn[threadIdx.x >> 6] = mn;
      s_mx[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < QK_WG / 64; w++) {
        mn = min(mn, s_mn[w]);
        mx = max(mx, s_mx[w]);
      }
      mn = min(mn, s_mn[0]);
      mx = max(mx, s_mx[0]);
      qk_gptr<double>(a.ptr[1])[z] = mn;
      qk_gptr<double>(a.ptr[2])[z] = mx;
    }
  }
}

/* ------------------------------------------------------------------ */
/* registry                                                             */
/* ------------------------------------------------------------------ */
static const QK_LibraryKernel g_library[] = {
    {"fill_u32", (const void*)k_fill_u32, QK_WG},
    {"bandwidth_read", (const void*)k_bandwidth_read, QK_WG},
    {"bandwidth_copy", (const void*)k_bandwidth_copy, QK_WG},
    {"aggregate_global", (const void*)k_aggregate_global, QK_WG},
    {"aggregate_global_or", (const void*)k_aggregate_global_or, QK_WG},
    {"aggregate_global_fold", (const void*)k_aggregate_global_fold, QK_WG},
    {"scan_simple", (const void*)k_scan_simple, QK_WG},
    {"scan_simple_or", (const void*)k_scan_simple_or, QK_WG},
    {"scan_filter", (const void*)k_scan_filter, QK_WG},
    {"aggregate_dense", (const void*)k_aggregate_dense, QK_WG},
    {"aggregate_dense_or", (const void*)k_aggregate_dense_or, QK_WG},
    {"aggregate_dense_wide", (const void*)k_aggregate_dense_wide, QK_WG},
    {"aggregate_dense_wide_or", (const void*)k_aggregate_dense_wide_or, QK_WG},
    {"aggregate_dense_narrow_s", (const void*)k_aggregate_dense_narrow_s, QK_NARROW_WG},
    {"aggregate_dense_narrow_m", (const void*)k_aggregate_dense_narrow_m, QK_NARROW_WG},
    {"aggregate_dense_narrow_sm", (const void*)k_aggregate_dense_narrow_sm, QK_NARROW_WG},
    {"aggregate_dense_narrow_s_or", (const void*)k_aggregate_dense_narrow_s_or, QK_NARROW_WG},
    {"aggregate_dense_narrow_m_or", (const void*)k_aggregate_dense_narrow_m_or, QK_NARROW_WG},
    {"aggregate_dense_narrow_sm_or", (const void*)k_aggregate_dense_narrow_sm_or, QK_NARROW_WG},
    {"join_build", (const void*)k_join_build, QK_WG},
    {"join_probe", (const void*)k_join_probe, QK_WG},
    {"aggregate_hash", (const void*)k_aggregate_hash, QK_WG},
    {"aggregate_hash_or", (const void*)k_aggregate_hash_or, QK_WG},
    {"shared_scan_256_1", (const void*)k_shared_scan_256_1, 256},
    {"shared_scan_256_2", (const void*)k_shared_scan_256_2, 256},
    {"shared_scan_256_4", (const void*)k_shared_scan_256_4, 256},
    {"shared_scan_512_1", (const void*)k_shared_scan_512_1, 512},
    {"shared_scan_512_2", (const void*)k_shared_scan_512_2, 512},
    {"shared_scan_1024_1", (const void*)k_shared_scan_1024_1, 1024},
    {"shared_scan_1024_1_r", (const void*)k_shared_scan_1024_1_r, 1024},
    {"shared_reduce", (const void*)k_shared_reduce, QK_WG},
    {"hash_compact", (const void*)k_hash_compact, QK_WG},
    {"join_part_build", (const void*)k_join_part_build, QK_WG},
    {"join_part_hist", (const void*)k_join_part_hist, QK_WG},
    {"join_part_offsets", (const void*)k_join_part_offsets, QK_WG},
    {"join_part_scatter", (const void*)k_join_part_scatter, QK_WG},
    {"join_part_scatter_0", (const void*)k_join_part_scatter_0, QK_WG},
    {"join_part_scatter_1", (const void*)k_join_part_scatter_1, QK_WG},
    {"join_part_scatter_2", (const void*)k_join_part_scatter_2, QK_WG},
    {"join_part_hist_512", (const void*)k_join_part_hist_512, 512},
    {"join_part_scatter_wc_0", (const void*)k_join_part_scatter_wc_0, QK_WC_WG},
    {"join_part_scatter_wc_1", (const void*)k_join_part_scatter_wc_1, QK_WC_WG},
    {"join_part_probe", (const void*)k_join_part_probe, QK_WG},
    {"join_part_probe_512", (const void*)k_join_part_probe_512, 512},
    {"join_part_probe_0", (const void*)k_join_part_probe_0, QK_WG},
    {"join_part_probe_0_512", (const void*)k_join_part_probe_0_512, 512},
    {"join_part_probe_1s", (const void*)k_join_part_probe_1s, QK_WG},
    {"join_part_probe_1s_512", (const void*)k_join_part_probe_1s_512, 512},
    {"join_part_probe_1m", (const void*)k_join_part_probe_1m, QK_WG},
    {"join_part_probe_1m_512", (const void*)k_join_part_probe_1m_512, 512},
    {"join_part_probe_1sm", (const void*)k_join_part_probe_1sm, QK_WG},
    {"join_part_probe_1sm_512", (const void*)k_join_part_probe_1sm_512, 512},
    {"join_part_probe_2s", (const void*)k_join_part_probe_2s, QK_WG},
    {"join_part_probe_2s_512", (const void*)k_join_part_probe_2s_512, 512},
    {"join_part_probe_2sm", (const void*)k_join_part_probe_2sm, QK_WG},
    {"join_part_probe_2sm_512", (const void*)k_join_part_probe_2sm_512, 512},
    {"join_part_probe_0_row", (const void*)k_join_part_probe_0_row, QK_WG},
    {"join_part_probe_0_row_512", (const void*)k_join_part_probe_0_row_512, 512},
    {"join_part_probe_1s_row", (const void*)k_join_part_probe_1s_row, QK_WG},
    {"join_part_probe_1s_row_512", (const void*)k_join_part_probe_1s_row_512, 512},
    {"join_part_probe_1sm_row", (const void*)k_join_part_probe_1sm_row, QK_WG},
    {"join_part_probe_1sm_row_512", (const void*)k_join_part_probe_1sm_row_512, 512},
    {"join_part_probe_pk_0_512", (const void*)k_join_part_probe_pk_0_512, 512},
    {"join_part_probe_pk_1s_512", (const void*)k_join_part_probe_pk_1s_512, 512},
    {"join_part_probe_pk_1m_512", (const void*)k_join_part_probe_pk_1m_512, 512},
    {"join_part_probe_pk_1sm_512", (const void*)k_join_part_probe_pk_1sm_512, 512},
    {"join_part_probe_pk_2s_512", (const void*)k_join_part_probe_pk_2s_512, 512},
    {"join_part_probe_pk_2sm_512", (const void*)k_join_part_probe_pk_2sm_512, 512},
    {"encode_column", (const void*)k_encode_column, QK_WG},
    {"seq_check", (const void*)k_seq_check, QK_WG},
    {"gather", (const void*)k_gather, QK_WG},
    {"scan_u32_blocks", (const void*)k_scan_u32_blocks, QK_WG},
    {"scan_u32_top", (const void*)k_scan_u32_top, QK_WG},
    {"scan_u32_add", (const void*)k_scan_u32_add, QK_WG},
    {"hjoin_build", (const void*)k_hjoin_build, QK_WG},
    {"hjoin_scatter", (const void*)k_hjoin_scatter, QK_WG},
    {"hjoin_sort_small", (const void*)k_hjoin_sort_small, QK_WG},
    {"hjoin_sort_big", (const void*)k_hjoin_sort_big, QK_WG},
    {"hjoin_probe", (const void*)k_hjoin_probe, QK_WG},
    {"gen_events", (const void*)k_gen_events, QK_WG},
    {"gen_scale_events", (const void*)k_gen_scale_events, QK_WG},
    {"gen_fact_fk", (const void*)k_gen_fact_fk, QK_WG},
    {"zonemap_build", (const void*)k_zonemap_build, QK_WG},
};

extern "C" const QK_LibraryKernel* qk_library_find(const char* name)
{
  for (unsigned i = 0; i < sizeof(g_library) / sizeof(g_library[0]); i++) {
    const char* a = g_library[i].name;
    const char* b = name;
    while (*a && *a == *b) {
      a++;
      b++;
    }
    if (*a == 0 && *b == 0) return &g_library[i];
  }
  return nullptr;
}

extern "C" U32 qk_library_count(void) { return (U32)(sizeof(g_library) / sizeof(g_library[0])); }

extern "C" const QK_LibraryKernel* qk_library_at(U32 index)
{
  return index < qk_library_count() ? &g_library[index] : nullptr;
}
