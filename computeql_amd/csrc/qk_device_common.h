/*
 * qk_device_common.h - device code shared by the library kernels
 * (qk_library.hip) and the per-query kernels compiled at run time with
 * hiprtc (query_kernel.c embeds this file verbatim as the prelude of every
 * generated source).  It must therefore include no headers.
 *
 * Scan model (gfx950, wave64): a thread owns a quad of 4 consecutive rows
 * read with one 16-byte load per 4-byte column; each block owns one
 * contiguous range of quads and walks it in steps of 256 threads x QK_UNROLL
 * quads, so the loads of a wave cover 1 KiB contiguous per column.  Row
 * matches are 4-bit masks per quad; stream compaction is wave64 ballot-free
 * prefix sums over the masks' popcounts, which keeps row ids in ascending
 * order (the reference's scan_simple emits them in arbitrary order,
 * scan_simple.comp:330-383).
 */
#ifndef QK_DEVICE_COMMON_H
#define QK_DEVICE_COMMON_H

#include "qk_abi.h"

typedef unsigned int qk_u32;
typedef unsigned long long qk_u64;
typedef int qk_s32;
typedef long long qk_s64;
typedef qk_u32 qk_u32x4 __attribute__((ext_vector_type(4)));
typedef qk_u64 qk_u64x2 __attribute__((ext_vector_type(2)));

#ifndef CQLHIP_H
/* same layout as QK_Args of cqlhip.h */
typedef struct QK_Args {
  qk_u64 ptr[16];
  qk_u64 bytes[16];
  qk_u32 words[QK_PUSH_WORD_COUNT];
  qk_u32 global_work_size;
  qk_u32 grid_blocks;
} QK_Args;
#endif

#define QK_WG 256
#define QK_UNROLL 4

/* column type codes (CQL_ColumnType, cql.h:78-95) */
#define QK_T_U32 1
#define QK_T_U64 2
#define QK_T_F32 3
#define QK_T_F64 4
#define QK_T_I32 7
#define QK_T_I64 8

/* scan leaf flags (query_exec.c:1372-1374) */
#define QK_LEAF_INT 0x10000u
#define QK_LEAF_NEVER 0x20000u
#define QK_LEAF_BIAS 0x40000u
#define QK_LEAF_NULL_TEST 0x80000u /* op 5 IS NULL, 6 IS NOT NULL: the null bitmap alone */
#define QK_LEAF_NEGATE 0x100000u   /* the comparison's complement (!=, NOT IN); NULL still fails */
#define QK_LEAF_SET 0x200000u      /* IN a set: (code - lo) < 64 and bit (code - lo) of words 2 | 4 << 32 */
/* a WHERE's leaves are followed by its terms (the leaves of each
 * conjunction of a disjunction, a byte each) at this word */
#define QK_TERM_WORD 48

/* scan bindings (query_exec.h:4-40) */
#define QK_B_LEAVES 0
#define QK_B_CONSTS 1
#define QK_B_BLOCK_COUNTS 2
#define QK_B_OUT_IDS 3
#define QK_B_OUT_COUNT 4
#define QK_B_SKIP 5
#define QK_B_COL0 6

/* scan push slots (U64): 0 row_count, 3 skip shift, 4 skip enabled,
 * 5 first row (chunk start), 6 mode, 7 leaf count */
#define QK_P_ROWS 0
#define QK_P_SKIP_SHIFT 3
#define QK_P_SKIP_ON 4
#define QK_P_ROW_BEGIN 5
#define QK_P_MODE 6
#define QK_P_LEAVES 7

/* modes 0 (count) and 1 (block counts) both leave one count per block in
 * binding 2; mode 2 writes the ordered row ids */
#define QK_MODE_COUNT 0
#define QK_MODE_BLOCK_COUNTS 1
#define QK_MODE_WRITE_IDS 2

__device__ __forceinline__ qk_u64 qk_push(const QK_Args& a, int slot)
{
  return (qk_u64)a.words[2 * slot] | ((qk_u64)a.words[2 * slot + 1] << 32);
}

/* Buffer addresses arrive as integers inside the by-value argument struct,
 * so on their own the compiler cannot tell that they point to global memory
 * and emits flat accesses, which also count against the LDS wait counter:
 * every wait for an LDS result would then wait for all loads in flight.
 * Making the pointer in address space 1 and casting it to a generic pointer
 * lets address-space inference turn the accesses through it into global_*
 * instructions, as it does for pointer kernel arguments. */
#define QK_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T* qk_gptr(qk_u64 addr)
{
  return (T*)(QK_GLOBAL T*)addr;
}

/* a load of column data the kernel reads once: non-temporal.  Measured on
 * gfx950 with 16-byte loads over a 4 GB column: 7.1-7.3 TB/s with the hint
 * against 6.2-6.6 TB/s without it, at every grid size tried. */
template <class T>
__device__ __forceinline__ T qk_ld(const T* p)
{
  return __builtin_nontemporal_load(p);
}

template <class T>
__device__ __forceinline__ const T* qk_ptr(const QK_Args& a, int binding)
{
  return qk_gptr<const T>(a.ptr[binding]);
}

template <class T>
__device__ __forceinline__ T* qk_wptr(const QK_Args& a, int binding)
{
  return qk_gptr<T>(a.ptr[binding]);
}

/* the contiguous quad range [b0,b1) of this block over [q0,q1), a multiple
 * of one step so no two blocks share a step */
__device__ __forceinline__ void qk_block_range(qk_u64 q0, qk_u64 q1, qk_u64& b0, qk_u64& b1)
{
  qk_u64 n = q1 - q0;
  qk_u64 step = (qk_u64)QK_WG * QK_UNROLL;
  qk_u64 per = (n + gridDim.x - 1) / gridDim.x;
  per = (per + step - 1) / step * step;
  b0 = q0 + per * blockIdx.x;
  if (b0 > q1) b0 = q1;
  b1 = b0 + per;
  if (b1 > q1) b1 = q1;
}

/* valid-row mask of quad q for rows [r0,r1) */
__device__ __forceinline__ qk_u32 qk_quad_valid(qk_u64 q, qk_u64 r0, qk_u64 r1)
{
  qk_u64 first = q * 4;
  qk_u32 m = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) m |= ((first + k >= r0) && (first + k < r1)) ? (1u << k) : 0u;
  return m;
}

/* zone-map skip bitmap: set bit = every row of the zone is skipped */
__device__ __forceinline__ bool qk_quad_skipped(const QK_Args& a, qk_u64 q)
{
  if (!qk_push(a, QK_P_SKIP_ON)) return false;
  qk_u64 zone = (q * 4) >> qk_push(a, QK_P_SKIP_SHIFT);
  const qk_u32* bits = qk_ptr<qk_u32>(a, QK_B_SKIP);
  return (bits[zone >> 5] >> (zone & 31)) & 1u;
}

/* four values of a column as f64 (exact for every 4-byte type) */
__device__ __forceinline__ void qk_load4_f64(const QK_Args& a, qk_u32 slot, qk_u32 type, qk_u64 q, double v[4])
{
  qk_u64 base = a.ptr[QK_B_COL0 + slot];
  if (type == QK_T_U32 || type == QK_T_I32 || type == QK_T_F32) {
    qk_u32x4 x = qk_ld(qk_gptr<const qk_u32x4>(base) + q);
    qk_u32 w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (type == QK_T_U32) v[k] = (double)w[k];
      else if (type == QK_T_I32) v[k] = (double)(qk_s32)w[k];
      else v[k] = (double)__uint_as_float(w[k]);
    }
  } else {
    const qk_u64x2* p = qk_gptr<const qk_u64x2>(base) + 2 * q;
    qk_u64x2 x0 = qk_ld(p), x1 = qk_ld(p + 1);
    qk_u64 w[4] = {x0.x, x0.y, x1.x, x1.y};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (type == QK_T_F64) v[k] = __longlong_as_double((qk_s64)w[k]);
      else if (type == QK_T_I64) v[k] = (double)(qk_s64)w[k];
      else v[k] = (double)w[k];
    }
  }
}

/* ---- narrow column encodings (qe_column_encoding / qe_encoding_shader_width,
 * query_exec.c:1555-1600): 0 full width, 1 u8 codes, 2 u16 codes, 3 a
 * sequence (the row number is the code, nothing stored).  A kernel given an
 * encoded column works on codes = value - base: integer leaves arrive with
 * their interval rewritten to codes, dense keys with key_min rewritten, and
 * the host adds the base back to sums, minima and maxima. ---- */
__device__ __forceinline__ qk_u32x4 qk_quad_codes(qk_u64 col, qk_u32 enc, qk_u64 q)
{
  if (enc == 0) return qk_ld(qk_gptr<const qk_u32x4>(col) + q);
  if (enc == 1) {
    const qk_u32 x = qk_ld(qk_gptr<const qk_u32>(col) + q);
    return qk_u32x4{x & 0xffu, (x >> 8) & 0xffu, (x >> 16) & 0xffu, x >> 24};
  }
  if (enc == 2) {
    const qk_u64 x = qk_ld(qk_gptr<const qk_u64>(col) + q);
    const qk_u32 lo = (qk_u32)x, hi = (qk_u32)(x >> 32);
    return qk_u32x4{lo & 0xffffu, lo >> 16, hi & 0xffffu, hi >> 16};
  }
  const qk_u32 r = (qk_u32)(q * 4);
  return qk_u32x4{r, r + 1u, r + 2u, r + 3u};
}

/* one numeric leaf over a quad (qe_simple_leaf_words layout,
 * query_exec.c:1442-1456; integer form :1408-1438) */
__device__ __forceinline__ qk_u32 qk_leaf_quad(const QK_Args& a, const qk_u32* w, qk_u64 q)
{
  qk_u32 slot = w[0], tw = w[1], op = w[2];
  qk_u32 type = tw & 0xffu;
  if (tw & QK_LEAF_NEVER) return 0u;
  qk_u32 m = 0;
  if ((tw & QK_LEAF_INT) && (type == QK_T_U32 || type == QK_T_I32)) {
    qk_u32 lo = w[3], span = w[4];
    qk_u32 bias = (tw & QK_LEAF_BIAS) ? 0x80000000u : 0u;
    qk_u32x4 x = qk_quad_codes(a.ptr[QK_B_COL0 + slot], (tw >> 8) & 0xffu, q);
    m |= (((x.x ^ bias) - lo) <= span) ? 1u : 0u;
    m |= (((x.y ^ bias) - lo) <= span) ? 2u : 0u;
    m |= (((x.z ^ bias) - lo) <= span) ? 4u : 0u;
    m |= (((x.w ^ bias) - lo) <= span) ? 8u : 0u;
    return m;
  }
  double t = __longlong_as_double((qk_s64)((qk_u64)w[3] | ((qk_u64)w[4] << 32)));
  double v[4];
  qk_load4_f64(a, slot, type, q, v);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    bool p;
    switch (op) {
      case 0: p = v[k] == t; break;
      case 1: p = v[k] < t; break;
      case 2: p = v[k] <= t; break;
      case 3: p = v[k] > t; break;
      default: p = v[k] >= t; break;
    }
    m |= p ? (1u << k) : 0u;
  }
  return m;
}

/* AND of leaf_count leaves (<= 8) over one quad */
__device__ __forceinline__ qk_u32 qk_leaves_quad(const QK_Args& a, const qk_u32* leaves, qk_u32 leaf_count, qk_u64 q,
                                                 qk_u32 mask)
{
  for (qk_u32 l = 0; l < leaf_count && l < 8u; l++) mask &= qk_leaf_quad(a, leaves + 6 * l, q);
  return mask;
}

/* the leaves of a WHERE held in registers for a whole kernel: loaded once
 * (uniform, scalar loads) instead of once per step.  `packed` is the leaf
 * count, with the term count in bits 8-15: 0 terms = the AND of the
 * leaves, else a row passes when every leaf of one of the terms passes */
struct QkLeafSet {
  qk_u32 n;
  qk_u32 terms;
  qk_u32 w[8][6];
  qk_u32 term_words[4];
};

template <bool TERMS = false>
__device__ __forceinline__ QkLeafSet qk_load_leaves(const qk_u32* leaves, qk_u32 packed)
{
  QkLeafSet L;
  const qk_u32 n = packed & 0xffu, terms = (packed >> 8) & 0xffu;
  L.n = n < 8u ? n : 8u;
  L.terms = TERMS ? (terms < 16u ? terms : 16u) : 0u;
#pragma unroll
  for (int l = 0; l < 8; l++)
#pragma unroll
    for (int i = 0; i < 6; i++)
      L.w[l][i] = __builtin_amdgcn_readfirstlane(((qk_u32)l < L.n) ? leaves[6 * l + i] : 0u);
#pragma unroll
  for (int i = 0; i < 4; i++)
    L.term_words[i] = TERMS ? __builtin_amdgcn_readfirstlane(L.terms ? leaves[QK_TERM_WORD + i] : 0u) : 0u;
  return L;
}

/* the rows of a quad passing a disjunction of conjunctions: p holds leaf
 * l's four row bits at bits 4l..4l+3.  A term's leaf byte is spread to a
 * nibble mask (uniform, scalar work); a row misses the term when one of
 * the term's leaves has the row's bit clear.  The term words are read
 * with static indexes, so the leaf set stays in registers. */
__device__ __forceinline__ qk_u32 qk_terms_quad(const QkLeafSet& L, qk_u32 p)
{
  qk_u32 out = 0u;
  const qk_u32 w0 = L.term_words[0], w1 = L.term_words[1], w2 = L.term_words[2], w3 = L.term_words[3];
  for (qk_u32 t = 0; t < L.terms; t++) {
    const qk_u32 word = t < 4u ? w0 : (t < 8u ? w1 : (t < 12u ? w2 : w3));
    qk_u32 n = (word >> (8u * (t & 3u))) & 0xffu;
    n = (n | (n << 12)) & 0x000f000fu;
    n = (n | (n << 6)) & 0x03030303u;
    n = (n | (n << 3)) & 0x11111111u;
    const qk_u32 mask = n * 0xfu;
    qk_u32 miss = ~p & mask;
    miss |= miss >> 16;
    miss |= miss >> 8;
    miss |= miss >> 4;
    out |= ~miss & 0xfu;
  }
  return out;
}

/* the WHERE over the QK_UNROLL quads of a step, leaf-major: the loads of
 * one leaf for every quad are issued together before any compare.  Each
 * leaf's row bits are ANDed into m, or with TERMS (a kernel form of its
 * own, so the AND form keeps its registers) gathered four bits a leaf and
 * matched against the terms */
template <bool TERMS = false>
__device__ __forceinline__ void qk_leaves_step(const QK_Args& a, const QkLeafSet& L, const qk_u64* q, qk_u32* m)
{
  qk_u32 p[QK_UNROLL];
  if (TERMS) {
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) p[j] = 0u;
  }
#pragma unroll
  for (int l = 0; l < 8; l++) {
    if ((qk_u32)l >= L.n) break;
    const qk_u32* w = L.w[l];
    qk_u32 slot = w[0], tw = w[1];
    qk_u32 type = tw & 0xffu;
    qk_u32 mm[QK_UNROLL];
    if (tw & QK_LEAF_NULL_TEST) {
      /* IS NULL keeps the rows whose bit is set, IS NOT NULL the others; a
       * column without NULLs (word 5 zero) has no bit set */
      const bool is_null = w[2] == 5u;
      if (w[5]) {
        const qk_u32* nb = qk_gptr<const qk_u32>(a.ptr[QK_B_COL0 + slot] + (qk_u64)w[5] * 16);
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) {
          const qk_u32 bits = (nb[q[j] >> 3] >> ((qk_u32)(q[j] & 7) * 4)) & 0xfu;
          mm[j] = is_null ? bits : (~bits & 0xfu);
        }
      } else {
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) mm[j] = is_null ? 0u : 0xfu;
      }
    } else {
      if (tw & QK_LEAF_NEVER) {
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) mm[j] = 0u;
      } else if ((tw & QK_LEAF_INT) && (type == QK_T_U32 || type == QK_T_I32)) {
        const qk_u32 lo = w[3];
        const qk_u32 bias = (tw & QK_LEAF_BIAS) ? 0x80000000u : 0u;
        const qk_u32 enc = (tw >> 8) & 0xffu;
        qk_u32x4 x[QK_UNROLL];
        if (enc == 0) {
          const qk_u32x4* col = qk_ptr<qk_u32x4>(a, QK_B_COL0 + slot);
#pragma unroll
          for (int j = 0; j < QK_UNROLL; j++) x[j] = qk_ld(col + q[j]);
        } else {
#pragma unroll
          for (int j = 0; j < QK_UNROLL; j++) x[j] = qk_quad_codes(a.ptr[QK_B_COL0 + slot], enc, q[j]);
        }
        if (tw & QK_LEAF_SET) {
          /* IN a set of values lo + i, i < 64 */
          const qk_u64 set = (qk_u64)w[2] | ((qk_u64)w[4] << 32);
#pragma unroll
          for (int j = 0; j < QK_UNROLL; j++) {
            const qk_u32 d0 = (x[j].x ^ bias) - lo, d1 = (x[j].y ^ bias) - lo;
            const qk_u32 d2 = (x[j].z ^ bias) - lo, d3 = (x[j].w ^ bias) - lo;
            qk_u32 r = 0;
            r |= (d0 < 64u && ((set >> d0) & 1u)) ? 1u : 0u;
            r |= (d1 < 64u && ((set >> d1) & 1u)) ? 2u : 0u;
            r |= (d2 < 64u && ((set >> d2) & 1u)) ? 4u : 0u;
            r |= (d3 < 64u && ((set >> d3) & 1u)) ? 8u : 0u;
            mm[j] = r;
          }
        } else {
          const qk_u32 span = w[4];
#pragma unroll
          for (int j = 0; j < QK_UNROLL; j++) {
            qk_u32 r = 0;
            r |= (((x[j].x ^ bias) - lo) <= span) ? 1u : 0u;
            r |= (((x[j].y ^ bias) - lo) <= span) ? 2u : 0u;
            r |= (((x[j].z ^ bias) - lo) <= span) ? 4u : 0u;
            r |= (((x[j].w ^ bias) - lo) <= span) ? 8u : 0u;
            mm[j] = r;
          }
        }
      } else {
        qk_u32 ww[6] = {w[0], w[1] & ~QK_LEAF_NEGATE, w[2], w[3], w[4], 0u};
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) mm[j] = qk_leaf_quad(a, ww, q[j]);
      }
      /* != and NOT IN take the complement */
      if (tw & QK_LEAF_NEGATE) {
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) mm[j] ^= 0xfu;
      }
      /* word 5: the column's null bitmap (16-byte units from its base); a
       * NULL row fails the leaf.  Quad q's four bits sit at bit
       * 4 (q mod 8) of word q / 8. */
      if (w[5]) {
        const qk_u32* nb = qk_gptr<const qk_u32>(a.ptr[QK_B_COL0 + slot] + (qk_u64)w[5] * 16);
#pragma unroll
        for (int j = 0; j < QK_UNROLL; j++) mm[j] &= ~((nb[q[j] >> 3] >> ((qk_u32)(q[j] & 7) * 4)) & 0xfu);
      }
    }
    if (TERMS) {
#pragma unroll
      for (int j = 0; j < QK_UNROLL; j++) p[j] |= (mm[j] & 0xfu) << (4 * l);
    } else {
#pragma unroll
      for (int j = 0; j < QK_UNROLL; j++) m[j] &= mm[j];
    }
  }
  if (TERMS) {
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) m[j] &= qk_terms_quad(L, p[j]);
  }
}

/* the QK_UNROLL quads a thread owns in the step starting at s: indices
 * clamped into [.., b1) so every load is in bounds, and valid-row masks
 * (0 for quads past the block range or in skipped zones) */
__device__ __forceinline__ void qk_step_quads(const QK_Args& a, qk_u64 s, qk_u64 b1, qk_u64 r0, qk_u64 r1, qk_u64* q,
                                              qk_u32* valid)
{
#pragma unroll
  for (int j = 0; j < QK_UNROLL; j++) {
    qk_u64 qq = s + (qk_u64)j * QK_WG + threadIdx.x;
    bool in = qq < b1;
    qk_u32 v = in ? qk_quad_valid(qq, r0, r1) : 0u;
    if (v && qk_quad_skipped(a, qq)) v = 0u;
    valid[j] = v;
    q[j] = in ? qq : b1 - 1;
  }
}

/* Block barrier that orders LDS accesses only.  __syncthreads() is also a
 * workgroup release fence, so it first waits for every global store of the
 * wave to complete; in a loop that stores results and then rebuilds LDS
 * state, that wait puts the HBM write latency on every iteration.  This
 * barrier waits for the wave's LDS operations (lgkmcnt) and nothing else;
 * the memory clobber keeps the compiler from moving accesses across it. */
__device__ __forceinline__ void qk_lds_barrier()
{
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

/* ---- wave64 / block reductions ---- */
__device__ __forceinline__ qk_u64 qk_wave_sum_u64(qk_u64 v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ qk_u32 qk_wave_incl_scan_u32(qk_u32 v)
{
  qk_u32 lane = __lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    qk_u32 y = __shfl_up(v, o, 64);
    if (lane >= (qk_u32)o) v += y;
  }
  return v;
}

/* block-wide sum of a u64, result valid in every thread */
template <int WG>
__device__ __forceinline__ qk_u64 qk_block_sum_u64_wg(qk_u64 v, qk_u64* lds)
{
  v = qk_wave_sum_u64(v);
  qk_u32 wave = threadIdx.x >> 6;
  __syncthreads();
  if (__lane_id() == 0) lds[wave] = v;
  __syncthreads();
  qk_u64 t = 0;
#pragma unroll
  for (int w = 0; w < WG / 64; w++) t += lds[w];
  return t;
}

__device__ __forceinline__ qk_u64 qk_block_sum_u64(qk_u64 v, qk_u64* lds4) { return qk_block_sum_u64_wg<QK_WG>(v, lds4); }

/* block exclusive scan of a u32 per thread over a WG-thread block; returns
 * the exclusive prefix and the block total; lds holds WG / 64 words */
template <int WG, bool LDS_ONLY>
__device__ __forceinline__ qk_u32 qk_block_excl_scan_wg(qk_u32 v, qk_u32* lds, qk_u32& total)
{
  qk_u32 incl = qk_wave_incl_scan_u32(v);
  qk_u32 wave = threadIdx.x >> 6;
  /* LDS_ONLY: the barriers order LDS only (qk_lds_barrier), for loops
   * whose global stores need no ordering against the scan */
  if (LDS_ONLY) qk_lds_barrier();
  else __syncthreads();
  if (__lane_id() == 63) lds[wave] = incl;
  if (LDS_ONLY) qk_lds_barrier();
  else __syncthreads();
  qk_u32 before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < WG / 64; w++) {
    qk_u32 x = lds[w];
    before += ((qk_u32)w < wave) ? x : 0u;
    all += x;
  }
  total = all;
  return before + incl - v;
}

/* the same over a QK_WG-thread block; lds4 holds QK_WG / 64 words */
template <bool LDS_ONLY = false>
__device__ __forceinline__ qk_u32 qk_block_excl_scan_u32(qk_u32 v, qk_u32* lds4, qk_u32& total)
{
  return qk_block_excl_scan_wg<QK_WG, LDS_ONLY>(v, lds4, total);
}

/* ---- the scan driver: count / per-block counts / ordered row ids ----
 * pred(q, mask) returns the match mask of quad q given the valid mask. */
/* a per-quad predicate p(q, valid) run over the quads of a step; the
 * quads' loads are unconditional (clamped indices) so they issue together */
template <class P>
struct QkPerQuad {
  P p;
  __device__ __forceinline__ void eval(const qk_u64* q, const qk_u32* valid, qk_u32* m) const
  {
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) m[j] = p(q[j], valid[j]) & valid[j];
  }
};

/* pred.eval(q[QK_UNROLL], valid[QK_UNROLL], m[QK_UNROLL]) */
template <class Pred>
__device__ __forceinline__ void qk_scan_driver_step(const QK_Args& a, Pred pred)
{
  __shared__ qk_u64 s_red[QK_WG / 64];
  __shared__ qk_u32 s_scan[QK_WG / 64];
  const qk_u64 rows = qk_push(a, QK_P_ROWS);
  const qk_u64 r0 = qk_push(a, QK_P_ROW_BEGIN);
  const qk_u64 r1 = r0 + rows;
  const qk_u32 mode = (qk_u32)qk_push(a, QK_P_MODE);
  qk_u64 b0, b1;
  qk_block_range(r0 / 4, (r1 + 3) / 4, b0, b1);

  if (mode != QK_MODE_WRITE_IDS) {
    qk_u64 count = 0;
    for (qk_u64 s = b0; s < b1; s += (qk_u64)QK_WG * QK_UNROLL) {
      qk_u64 q[QK_UNROLL];
      qk_u32 valid[QK_UNROLL], m[QK_UNROLL];
      qk_step_quads(a, s, b1, r0, r1, q, valid);
      pred.eval(q, valid, m);
#pragma unroll
      for (int j = 0; j < QK_UNROLL; j++) count += (qk_u64)__popc(m[j]);
    }
    /* count and block-count modes: one count per block, added up by the
     * caller.  (A last-block fold inside the kernel would need an agent-
     * scope release/acquire per block, which on this part writes back and
     * invalidates the XCD's L2 under the blocks still streaming.) */
    qk_u64 total = qk_block_sum_u64(count, s_red);
    qk_u64* bc = qk_wptr<qk_u64>(a, QK_B_BLOCK_COUNTS);
    if (threadIdx.x == 0) bc[blockIdx.x] = total;
    return;
  }

  /* ordered write: this block's first output index is the sum of the
   * counts of the blocks before it */
  const qk_u64* bc = qk_ptr<qk_u64>(a, QK_B_BLOCK_COUNTS);
  qk_u64 part = 0;
  for (qk_u32 b = threadIdx.x; b < blockIdx.x; b += QK_WG) part += bc[b];
  qk_u64 base = qk_block_sum_u64(part, s_red);
  qk_u64* out = qk_wptr<qk_u64>(a, QK_B_OUT_IDS);
  const qk_u64 cap = a.bytes[QK_B_OUT_IDS] / 8;
  for (qk_u64 s = b0; s < b1; s += (qk_u64)QK_WG * QK_UNROLL) {
    qk_u64 qs[QK_UNROLL];
    qk_u32 valid[QK_UNROLL], m[QK_UNROLL];
    qk_step_quads(a, s, b1, r0, r1, qs, valid);
    pred.eval(qs, valid, m);
#pragma unroll
    for (int j = 0; j < QK_UNROLL; j++) {
      qk_u32 total = 0;
      qk_u32 pre = qk_block_excl_scan_u32((qk_u32)__popc(m[j]), s_scan, total);
      qk_u64 q = qs[j];
      qk_u64 o = base + pre;
      qk_u32 mm = m[j];
      while (mm) {
        int k = __ffs(mm) - 1;
        if (o < cap) out[o] = q * 4 + (qk_u64)k;
        o++;
        mm &= mm - 1;
      }
      base += total;
    }
  }
}

/* per-quad predicates (generated kernels, the interpreter) */
template <class P>
__device__ __forceinline__ void qk_scan_driver(const QK_Args& a, P pred)
{
  qk_scan_driver_step(a, QkPerQuad<P>{pred});
}

#endif
