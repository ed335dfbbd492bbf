/*
 * query_kernel.c - per-query WHERE kernels generated as HIP source and
 * compiled at run time with hiprtc (the reference generates GLSL and
 * compiles it with shaderc: query_kernel.c:853-889 simple kernels,
 * :970-1673 qk_scan_source_build, :2048/:2072 entry points, :214/:252-276
 * two-hash naming, :367-452 the on-disk cache).
 *
 * The generator symbolically executes the WHERE bytecode (the same postfix
 * program the interpreter runs) into one boolean expression per row.
 * Literals stay in a device buffer (binding 1), so the kernel depends only
 * on the shape of the WHERE clause and queries that differ in literals share
 * one kernel.  Every generated source starts with qk_device_common.h and
 * reuses its scan driver, so generated and library kernels share the
 * launch ABI, the quad walk and the ordered compaction.
 *
 * Compiled code objects are cached in-process (the backend's LRU of loaded
 * modules) and on disk under $CQLHIP_KERNEL_CACHE (default
 * ~/.cache/cqlhip), keyed by two independent hashes of the source.
 */
#define _GNU_SOURCE
#include <hip/hiprtc.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdlib.h>
#include <sys/stat.h>
#include <unistd.h>

#include "qe_internal.h"

static const char qk_prelude[] =
#include "qk_prelude.inc"
    ;

typedef struct QK_Src {
  char* buf;
  size_t len, cap;
} QK_Src;

static void src_printf(QK_Src* s, const char* fmt, ...)
{
  va_list ap;
  for (;;) {
    size_t room = s->cap - s->len;
    va_start(ap, fmt);
    int n = vsnprintf(s->buf + s->len, room, fmt, ap);
    va_end(ap);
    if (n < 0) return;
    if ((size_t)n < room) {
      s->len += (size_t)n;
      return;
    }
    size_t ncap = (s->cap ? s->cap * 2 : 4096) + (size_t)n;
    char* nb = (char*)realloc(s->buf, ncap);
    if (!nb) return;
    s->buf = nb;
    s->cap = ncap;
  }
}

static U64 hash_fnv1a(const char* s, size_t n)
{
  U64 h = 14695981039346656037ull;
  for (size_t i = 0; i < n; i++) {
    h ^= (U8)s[i];
    h *= 1099511628211ull;
  }
  return h;
}

static U32 hash_djb2(const char* s, size_t n)
{
  U32 h = 5381;
  for (size_t i = 0; i < n; i++) h = h * 33u + (U8)s[i];
  return h;
}

static const char* cache_dir(char* buf, size_t n)
{
  const char* d = getenv("CQLHIP_KERNEL_CACHE");
  if (d && d[0]) return d;
  const char* home = getenv("HOME");
  snprintf(buf, n, "%s/.cache/cqlhip", home ? home : "/tmp");
  return buf;
}

static B32 cache_load(const char* name, char** out, size_t* out_size)
{
  char dir[512], path[768];
  snprintf(path, sizeof(path), "%s/%s.co", cache_dir(dir, sizeof(dir)), name);
  FILE* f = fopen(path, "rb");
  if (!f) return 0;
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  if (sz <= 0) {
    fclose(f);
    return 0;
  }
  char* b = (char*)malloc((size_t)sz);
  if (!b || fread(b, 1, (size_t)sz, f) != (size_t)sz) {
    free(b);
    fclose(f);
    return 0;
  }
  fclose(f);
  *out = b;
  *out_size = (size_t)sz;
  return 1;
}

static void cache_store(const char* name, const char* code, size_t size)
{
  if (qe_settings_u64("QE_KERNEL_DISK_CACHE", 1) == 0) return;
  char dir[512], path[768], tmp[800];
  const char* d = cache_dir(dir, sizeof(dir));
  char parent[512];
  snprintf(parent, sizeof(parent), "%s", d);
  char* slash = strrchr(parent, '/');
  if (slash && slash != parent) {
    *slash = 0;
    mkdir(parent, 0755);
  }
  mkdir(d, 0755);
  snprintf(path, sizeof(path), "%s/%s.co", d, name);
  snprintf(tmp, sizeof(tmp), "%s.%d.tmp", path, (int)getpid());
  FILE* f = fopen(tmp, "wb");
  if (!f) return;
  size_t w = fwrite(code, 1, size, f);
  fclose(f);
  if (w == size) rename(tmp, path);
  else remove(tmp);
}

/* compile (or find) the kernel for `src`; names carry two hashes */
static GPU_Kernel* qk_kernel_from_source(const char* prefix, QK_Src* src, U64* compile_ns, char* out_name)
{
  char name[64];
  snprintf(name, sizeof(name), "%s_%016llx_%08x", prefix, (unsigned long long)hash_fnv1a(src->buf, src->len),
           hash_djb2(src->buf, src->len));
  if (out_name) snprintf(out_name, 64, "%s", name);
  String8 empty = {0, 0};
  GPU_Kernel* k = gpu_kernel_alloc_from_spirv(qe_str8(name), empty);
  if (k) return k;
  U64 t0 = qe_now_ns();
  char* code = 0;
  size_t code_size = 0;
  if (!(qe_settings_u64("QE_KERNEL_DISK_CACHE", 1) && cache_load(name, &code, &code_size))) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src->buf, "qk_query.hip", 0, 0, 0) != HIPRTC_SUCCESS) {
      qe_set_error("hiprtcCreateProgram failed");
      return 0;
    }
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (r != HIPRTC_SUCCESS) {
      size_t log_size = 0;
      hiprtcGetProgramLogSize(prog, &log_size);
      char* log = (char*)malloc(log_size + 1);
      if (log) {
        hiprtcGetProgramLog(prog, log);
        log[log_size] = 0;
        qe_set_error("hiprtc: %.400s", log);
        free(log);
      }
      hiprtcDestroyProgram(&prog);
      return 0;
    }
    hiprtcGetCodeSize(prog, &code_size);
    code = (char*)malloc(code_size ? code_size : 1);
    if (!code || hiprtcGetCode(prog, code) != HIPRTC_SUCCESS) {
      free(code);
      hiprtcDestroyProgram(&prog);
      qe_set_error("hiprtcGetCode failed");
      return 0;
    }
    hiprtcDestroyProgram(&prog);
    cache_store(name, code, code_size);
  }
  String8 blob = {(U8*)code, code_size};
  k = gpu_kernel_alloc_from_spirv(qe_str8(name), blob);
  free(code);
  if (compile_ns) *compile_ns += qe_now_ns() - t0;
  if (!k) qe_set_error("loading generated kernel %s: %s", name, gpu_hip_last_error());
  return k;
}

/* ---------------- WHERE shapes ---------------- */
/* A generated kernel depends only on the bytecode words (operators, column
 * slots with their types, constant indices): literals travel in a buffer.
 * Shapes seen before map to their kernel's name, so a repeated shape skips
 * generating and hashing its source; the words themselves are kept and
 * compared, so two shapes never share an entry. */
#define QK_SHAPES 512
typedef struct QK_Shape {
  U64 hash;
  U32 word_count;
  U32* words;
  char name[64];  /* "" until its kernel is built */
  B32 seen;       /* run once (QE_ROUTE_CONJUNCTIONS) */
} QK_Shape;
static QK_Shape g_shapes[QK_SHAPES];
static pthread_mutex_t g_shape_lock = PTHREAD_MUTEX_INITIALIZER;

static U64 qk_shape_hash(const QE_Bytecode* p)
{
  return hash_fnv1a((const char*)p->words, p->word_count * sizeof(U32)) ^ ((U64)p->word_count << 52);
}

/* the entry of a shape (inserted when absent, evicting the slot's old
 * shape); callers hold g_shape_lock */
static QK_Shape* qk_shape_locked(const QE_Bytecode* p)
{
  const U64 h = qk_shape_hash(p);
  QK_Shape* e = &g_shapes[h % QK_SHAPES];
  if (e->words && e->hash == h && e->word_count == p->word_count &&
      !memcmp(e->words, p->words, p->word_count * sizeof(U32)))
    return e;
  U32* w = (U32*)malloc(p->word_count * sizeof(U32) + 4);
  if (!w) return 0;
  memcpy(w, p->words, p->word_count * sizeof(U32));
  free(e->words);
  memset(e, 0, sizeof(*e));
  e->hash = h;
  e->word_count = p->word_count;
  e->words = w;
  return e;
}

/* QE_ROUTE_CONJUNCTIONS (query_kernel.c:1735-1760): with it set, the first
 * run of a WHERE shape is left to the library kernels and its kernel is
 * generated when the shape comes back; 1 = run this one on the library */
B32 qk_route_defer(const QE_Bytecode* p)
{
  if (qe_settings_u64("QE_ROUTE_CONJUNCTIONS", 0) == 0) return 0;
  pthread_mutex_lock(&g_shape_lock);
  QK_Shape* e = qk_shape_locked(p);
  B32 defer = e && !e->seen && !e->name[0];
  if (e) e->seen = 1;
  pthread_mutex_unlock(&g_shape_lock);
  return defer;
}

/* ---------------- qk_scan_kernel: any AND/OR tree ---------------- */
enum {
  OP_PUSH_TRUE = 0,
  OP_LOAD_NUM_COL = 1,
  OP_PUSH_CONST = 2,
  OP_CMP_EQ = 3,
  OP_CMP_NE = 4,
  OP_CMP_LT = 5,
  OP_CMP_GT = 6,
  OP_CMP_LE = 7,
  OP_CMP_GE = 8,
  OP_AND = 9,
  OP_OR = 10,
  OP_HALT = 13,
  OP_PUSH_FALSE = 14,
};

#define QK_EXPR_MAX 8192
typedef struct {
  char* text;
  B32 is_bool;
} QK_Expr;

static char* qk_strdupf(const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  char* out = 0;
  if (vasprintf(&out, fmt, ap) < 0) out = 0;
  va_end(ap);
  return out;
}

static char* qk_as_bool(QK_Expr e) { return e.is_bool ? qk_strdupf("%s", e.text) : qk_strdupf("(%s != 0.0)", e.text); }
static char* qk_as_num(QK_Expr e) { return e.is_bool ? qk_strdupf("(%s ? 1.0 : 0.0)", e.text) : qk_strdupf("%s", e.text); }

GPU_Kernel* qk_scan_kernel(QE_Table* table, const QE_IRNode* nodes, U32 count, QE_ColumnSlots* slots, F64* consts,
                           U32* const_count, U64* compile_ns)
{
  static __thread QE_Bytecode prog;
  if (!qe_bytecode_program_build(table, nodes, count, &prog)) {
    qe_set_error("qk_scan_kernel: %s", prog.error);
    return 0;
  }
  /* a shape built before: its kernel by name, no source generated */
  char known[64] = "";
  pthread_mutex_lock(&g_shape_lock);
  {
    QK_Shape* e = qk_shape_locked(&prog);
    if (e) snprintf(known, sizeof(known), "%s", e->name);
  }
  pthread_mutex_unlock(&g_shape_lock);
  if (known[0]) {
    String8 empty = {0, 0};
    GPU_Kernel* k = gpu_kernel_alloc_from_spirv(qe_str8(known), empty);
    if (k) {
      *slots = prog.slots;
      memcpy(consts, prog.consts, sizeof(F64) * prog.const_count);
      *const_count = prog.const_count;
      return k;
    }
  }
  QK_Expr stack[64];
  int sp = 0;
  B32 ok = 1;
  U32 slot_type[QE_MAX_COLUMN_BINDINGS];
  U32 ip = 0;
  while (ip < prog.word_count && ok) {
    U32 op = prog.words[ip++];
    if (op == OP_HALT) break;
    if (sp >= 62) {
      ok = 0;
      break;
    }
    switch (op) {
      case OP_PUSH_TRUE: stack[sp++] = (QK_Expr){qk_strdupf("true"), 1}; break;
      case OP_PUSH_FALSE: stack[sp++] = (QK_Expr){qk_strdupf("false"), 1}; break;
      case OP_LOAD_NUM_COL: {
        U32 operand = prog.words[ip++];
        U32 slot = operand & 0xffu;
        slot_type[slot] = operand >> 8;
        stack[sp++] = (QK_Expr){qk_strdupf("v%u[k]", slot), 0};
      } break;
      case OP_PUSH_CONST: {
        U32 ci = prog.words[ip++];
        stack[sp++] = (QK_Expr){qk_strdupf("c%u", ci), 0};
      } break;
      default: {
        if (sp < 2) {
          ok = 0;
          break;
        }
        QK_Expr r = stack[--sp], l = stack[--sp];
        char* t = 0;
        if (op == OP_AND || op == OP_OR) {
          char* lb = qk_as_bool(l);
          char* rb = qk_as_bool(r);
          t = qk_strdupf("(%s %s %s)", lb, op == OP_AND ? "&&" : "||", rb);
          free(lb);
          free(rb);
        } else {
          const char* sym = op == OP_CMP_EQ ? "==" : op == OP_CMP_NE ? "!=" : op == OP_CMP_LT ? "<" :
                            op == OP_CMP_GT ? ">" : op == OP_CMP_LE ? "<=" : ">=";
          char* ln = qk_as_num(l);
          char* rn = qk_as_num(r);
          t = qk_strdupf("(%s %s %s)", ln, sym, rn);
          free(ln);
          free(rn);
        }
        free(l.text);
        free(r.text);
        if (!t || strlen(t) > QK_EXPR_MAX) ok = 0;
        stack[sp++] = (QK_Expr){t, 1};
      } break;
    }
  }
  if (!ok || sp != 1) {
    for (int i = 0; i < sp; i++) free(stack[i].text);
    qe_set_error("qk_scan_kernel: cannot generate this WHERE shape");
    return 0;
  }
  char* root = qk_as_bool(stack[0]);
  free(stack[0].text);

  QK_Src s = {0, 0, 0};
  src_printf(&s, "%s\n", qk_prelude);
  src_printf(&s, "struct QkPred {\n  const QK_Args* a;\n");
  src_printf(&s, "  __device__ __forceinline__ qk_u32 operator()(qk_u64 q, qk_u32 valid) const {\n");
  src_printf(&s, "    const double* cp = (const double*)a->ptr[QK_B_CONSTS];\n");
  for (U32 c = 0; c < prog.const_count; c++) src_printf(&s, "    const double c%u = cp[%u];\n", c, c);
  for (U32 sl = 0; sl < prog.slots.count; sl++)
    src_printf(&s, "    double v%u[4]; qk_load4_f64(*a, %u, %u, q, v%u);\n", sl, sl, slot_type[sl], sl);
  src_printf(&s, "    qk_u32 m = 0;\n#pragma unroll\n    for (int k = 0; k < 4; k++) m |= %s ? (1u << k) : 0u;\n", root);
  src_printf(&s, "    return m & valid;\n  }\n};\n");
  src_printf(&s, "extern \"C\" __global__ void __launch_bounds__(QK_WG) qk_main(QK_Args a) {\n");
  src_printf(&s, "  QkPred p{&a};\n  qk_scan_driver(a, p);\n}\n");
  free(root);
  char name[64] = "";
  GPU_Kernel* k = s.buf ? qk_kernel_from_source("qk_scan", &s, compile_ns, name) : 0;
  free(s.buf);
  if (!k) return 0;
  pthread_mutex_lock(&g_shape_lock);
  {
    QK_Shape* e = qk_shape_locked(&prog);
    if (e) snprintf(e->name, sizeof(e->name), "%s", name);
  }
  pthread_mutex_unlock(&g_shape_lock);
  *slots = prog.slots;
  memcpy(consts, prog.consts, sizeof(F64) * prog.const_count);
  *const_count = prog.const_count;
  return k;
}

/* ---------------- qk_simple_kernel: numeric AND leaves baked in ---------------- */
GPU_Kernel* qk_simple_kernel(QE_Table* table, const QE_Leaf* leaves, U32 leaf_count, QE_ColumnSlots* slots,
                             U64* compile_ns)
{
  QE_ColumnSlots sl;
  memset(&sl, 0, sizeof(sl));
  U32 leaf_slot[QE_SCAN_MAX_LEAVES];
  for (U32 l = 0; l < leaf_count; l++) leaf_slot[l] = qe_slot_of(&sl, leaves[l].column);
  B32 int_compare = qe_settings_u64("QE_SCAN_INT_COMPARE", 1) != 0;
  QK_Src s = {0, 0, 0};
  src_printf(&s, "%s\n", qk_prelude);
  src_printf(&s, "struct QkPred {\n  const QK_Args* a;\n");
  src_printf(&s, "  __device__ __forceinline__ qk_u32 operator()(qk_u64 q, qk_u32 valid) const {\n");
  src_printf(&s, "    const qk_u32* L = qk_ptr<qk_u32>(*a, QK_B_LEAVES);\n    qk_u32 m = valid;\n");
  for (U32 slot = 0; slot < sl.count; slot++) {
    U32 type = table->columns[sl.column[slot]].type;
    B32 as_int = int_compare && (type == CQL_ColumnType_U32 || type == CQL_ColumnType_I32);
    if (as_int) src_printf(&s, "    { qk_u32x4 x = ((const qk_u32x4*)a->ptr[QK_B_COL0 + %u])[q];\n", slot);
    else src_printf(&s, "    { double v[4]; qk_load4_f64(*a, %u, %u, q, v);\n", slot, type);
    for (U32 l = 0; l < leaf_count; l++) {
      if (leaf_slot[l] != slot) continue;
      if (as_int) {
        const char* bias = type == CQL_ColumnType_I32 ? " ^ 0x80000000u" : "";
        src_printf(&s,
                   "      { qk_u32 lo = L[%u], span = L[%u]; qk_u32 mm = 0;\n"
                   "        mm |= (((x.x%s) - lo) <= span) ? 1u : 0u; mm |= (((x.y%s) - lo) <= span) ? 2u : 0u;\n"
                   "        mm |= (((x.z%s) - lo) <= span) ? 4u : 0u; mm |= (((x.w%s) - lo) <= span) ? 8u : 0u;\n"
                   "        if (L[%u] & QK_LEAF_NEVER) mm = 0u; m &= mm; }\n",
                   6 * l + 3, 6 * l + 4, bias, bias, bias, bias, 6 * l + 1);
      } else {
        static const char* sym[5] = {"==", "<", "<=", ">", ">="};
        src_printf(&s,
                   "      { double t = __longlong_as_double((qk_s64)((qk_u64)L[%u] | ((qk_u64)L[%u] << 32))); qk_u32 mm = 0;\n"
                   "#pragma unroll\n        for (int k = 0; k < 4; k++) mm |= (v[k] %s t) ? (1u << k) : 0u;\n"
                   "        m &= mm; }\n",
                   6 * l + 3, 6 * l + 4, sym[leaves[l].op]);
      }
    }
    src_printf(&s, "    }\n");
  }
  src_printf(&s, "    return m;\n  }\n};\n");
  src_printf(&s, "extern \"C\" __global__ void __launch_bounds__(QK_WG) qk_main(QK_Args a) {\n");
  src_printf(&s, "  QkPred p{&a};\n  qk_scan_driver(a, p);\n}\n");
  GPU_Kernel* k = s.buf ? qk_kernel_from_source("qk_simple", &s, compile_ns, 0) : 0;
  free(s.buf);
  if (k) *slots = sl;
  return k;
}
