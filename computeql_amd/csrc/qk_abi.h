/*
 * qk_abi.h - push-word layout of the library kernels, shared by the host
 * operators (C) and the device code.  Plain macros only.
 *
 * Words 0..15 are the reference's 8 x U64 push constants
 * (gpu_vulkan.h:6-9); the scan kernels use U64 slots 0 (rows), 3 (zone
 * shift), 4 (skip on), 5 (first row), 6 (mode), 7 (leaf count), i.e. words
 * 0-1 and 6-15.  Operator parameters of the aggregate and join kernels live
 * in words 16..31, which no U64 slot overlaps.
 */
#ifndef QK_ABI_H
#define QK_ABI_H

#define QK_PUSH_WORD_COUNT 32

#define QK_W_ARGS 16     /* arg count | arg slots << 8 (4 bits per arg) */
#define QK_W_TYPES 17    /* arg column types, 8 bits per arg */
#define QK_W_KEY 18      /* key slot | key type << 8 | flags << 16 */
#define QK_W_KEY_MIN 19  /* code of the smallest key / group */
#define QK_W_RANGE 20    /* groups R */
#define QK_W_LOG2 21     /* log2 hash-table slots */
#define QK_W_PMODE 22    /* join_build payload mode */
#define QK_W_GMIN 23     /* join_build: group code minimum */
#define QK_W_GTYPE 24    /* join_build: group column type */
#define QK_W_ZTYPE 25    /* zonemap_build: column type */
#define QK_W_ZSHIFT 26   /* zonemap_build: log2 rows per zone */
#define QK_W_PBITS 27    /* partitioned join: log2 partitions */
#define QK_W_SBITS 28    /* partitioned join: log2 slots per partition table */
#define QK_W_ROWW 29     /* partitioned join: u32 words per partitioned row (1 + args) */
#define QK_W_SPLIT 30    /* partitioned join: probe blocks per partition */

#endif
