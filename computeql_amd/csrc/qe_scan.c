/*
 * qe_scan.c - WHERE compilation and the GPU scan driver.
 *
 *   qe_bytecode_program_build   query_exec.c:524-545 (+ qe_compile_condition :203-522)
 *   qe_collect_and_leaves       qe_cpu_scan_collect_leaves :1189-1235
 *   qe_scan_int_interval        :1378-1404
 *   qe_simple_leaf_words        :1442-1456 (+ qe_scan_int_leaf_words :1408-1438)
 *   qe_zonemap_prune            qe_scan_build_dispatch_ranges :13307-13545,
 *                               qe_zonemap_chunk_is_prunable :13294-13304
 *   qe_scan_filter              qe_scan_filter_selected :2495-3273
 *
 * The columns are resident in HBM whole, so the 16M-row chunk grid and the
 * slot pipeline of qe_scan_filter_streamed (:2242-2491) reduce to one
 * dispatch over the zone-map dispatch range.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>

#include "qe_internal.h"

/* QE_Opcode, query_exec.h:58-82 */
enum {
  QE_Opcode_PushTrue = 0,
  QE_Opcode_LoadNumCol = 1,
  QE_Opcode_PushConst = 2,
  QE_Opcode_CmpEq = 3,
  QE_Opcode_CmpNe = 4,
  QE_Opcode_CmpLt = 5,
  QE_Opcode_CmpGt = 6,
  QE_Opcode_CmpLe = 7,
  QE_Opcode_CmpGe = 8,
  QE_Opcode_And = 9,
  QE_Opcode_Or = 10,
  QE_Opcode_Halt = 13,
  QE_Opcode_PushFalse = 14,
};

U32 qe_ir_subtree_end(const QE_IRNode* nodes, U32 i)
{
  U32 end = i + 1;
  for (U32 c = 0; c < nodes[i].child_count; c++) end = qe_ir_subtree_end(nodes, end);
  return end;
}

U32 qe_slot_of(QE_ColumnSlots* slots, U32 column)
{
  for (U32 i = 0; i < slots->count; i++)
    if (slots->column[i] == column) return i;
  if (slots->count >= QE_MAX_COLUMN_BINDINGS) return ~0u;
  slots->column[slots->count] = column;
  return slots->count++;
}

static void qe_emit(QE_Bytecode* p, U32 w)
{
  if (p->word_count >= sizeof(p->words) / sizeof(p->words[0])) {
    p->overflow = 1;
    return;
  }
  p->words[p->word_count++] = w;
}

static U32 qe_add_const(QE_Bytecode* p, F64 v)
{
  if (p->const_count >= sizeof(p->consts) / sizeof(p->consts[0])) {
    p->overflow = 1;
    return 0;
  }
  p->consts[p->const_count] = v;
  return p->const_count++;
}

static U32 qe_cmp_opcode(U32 irop)
{
  switch (irop) {
    case QE_IROP_EQ: return QE_Opcode_CmpEq;
    case QE_IROP_NE: return QE_Opcode_CmpNe;
    case QE_IROP_LT: return QE_Opcode_CmpLt;
    case QE_IROP_GT: return QE_Opcode_CmpGt;
    case QE_IROP_LE: return QE_Opcode_CmpLe;
    case QE_IROP_GE: return QE_Opcode_CmpGe;
    default: return ~0u;
  }
}

/* qe_compile_load_value, query_exec.c:180-200 */
static B32 qe_compile_load_value(QE_Table* t, const QE_IRNode* n, QE_Bytecode* p)
{
  if (n->type == QE_IR_COLUMN) {
    if (n->column >= t->column_count) {
      snprintf(p->error, sizeof(p->error), "unknown column %u", n->column);
      return 0;
    }
    U32 slot = qe_slot_of(&p->slots, n->column);
    if (slot == ~0u) {
      snprintf(p->error, sizeof(p->error), "more than %d columns in WHERE", QE_MAX_COLUMN_BINDINGS);
      return 0;
    }
    qe_emit(p, QE_Opcode_LoadNumCol);
    qe_emit(p, (t->columns[n->column].type << 8) | (slot & 0xffu));
    return 1;
  }
  if (n->type == QE_IR_NUMERIC) {
    U32 ci = qe_add_const(p, n->value);
    qe_emit(p, QE_Opcode_PushConst);
    qe_emit(p, ci);
    return 1;
  }
  snprintf(p->error, sizeof(p->error), "unsupported operand node type %u", n->type);
  return 0;
}

/* qe_compile_condition, query_exec.c:203-522 (numeric subset) */
static B32 qe_compile_condition(QE_Table* t, const QE_IRNode* nodes, U32 i, QE_Bytecode* p)
{
  const QE_IRNode* n = &nodes[i];
  if (n->type != QE_IR_OPERATOR) return qe_compile_load_value(t, n, p); /* bare value: nonzero = true */
  U32 left = i + 1;
  U32 right = n->child_count > 1 ? qe_ir_subtree_end(nodes, left) : 0;
  switch (n->op) {
    case QE_IROP_AND:
    case QE_IROP_OR:
      if (n->child_count != 2) break;
      if (!qe_compile_condition(t, nodes, left, p) || !qe_compile_condition(t, nodes, right, p)) return 0;
      qe_emit(p, n->op == QE_IROP_AND ? QE_Opcode_And : QE_Opcode_Or);
      return 1;
    case QE_IROP_IS_NULL:
    case QE_IROP_IS_NOT_NULL:
      /* :233-249: columns on this path hold no NULLs */
      if (n->op == QE_IROP_IS_NOT_NULL) qe_emit(p, QE_Opcode_PushTrue);
      else {
        qe_emit(p, QE_Opcode_PushConst);
        qe_emit(p, qe_add_const(p, 0.0));
      }
      return 1;
    case QE_IROP_IN:
    case QE_IROP_NOT_IN: {
      /* :251-305: an OR chain of '=' (AND of '!=' for NOT IN), at most
       * QE_IN_LIST_GPU_MAX_ITEMS items */
      B32 is_not = n->op == QE_IROP_NOT_IN;
      if (n->child_count != 2 || nodes[left].type != QE_IR_COLUMN || nodes[right].type != QE_IR_INLIST) {
        snprintf(p->error, sizeof(p->error), "IN needs a column and a value list");
        return 0;
      }
      U32 items = nodes[right].child_count;
      if (items == 0) {
        qe_emit(p, is_not ? QE_Opcode_PushTrue : QE_Opcode_PushFalse);
        return 1;
      }
      if (items > qe_settings_u64("QE_IN_LIST_GPU_MAX_ITEMS", 32)) {
        snprintf(p->error, sizeof(p->error), "IN list too long for GPU bytecode");
        return 0;
      }
      for (U32 k = 0; k < items; k++) {
        if (!qe_compile_load_value(t, &nodes[left], p)) return 0;
        if (!qe_compile_load_value(t, &nodes[right + 1 + k], p)) return 0;
        qe_emit(p, is_not ? QE_Opcode_CmpNe : QE_Opcode_CmpEq);
        if (k > 0) qe_emit(p, is_not ? QE_Opcode_And : QE_Opcode_Or);
      }
      return 1;
    }
    default: {
      U32 opc = qe_cmp_opcode(n->op);
      if (opc == ~0u || n->child_count != 2) break;
      if (!qe_compile_load_value(t, &nodes[left], p) || !qe_compile_load_value(t, &nodes[right], p)) return 0;
      qe_emit(p, opc);
      return 1;
    }
  }
  snprintf(p->error, sizeof(p->error), "unsupported operator %u", n->op);
  return 0;
}

/* qe_bytecode_program_max_stack_depth, query_exec.c:547-600 */
static U32 qe_bytecode_max_stack_depth(const QE_Bytecode* p)
{
  S32 depth = 0, max_depth = 0;
  for (U32 ip = 0; ip < p->word_count;) {
    U32 op = p->words[ip++];
    if (op == QE_Opcode_LoadNumCol || op == QE_Opcode_PushConst) {
      ip++;
      depth++;
    } else if (op == QE_Opcode_PushTrue || op == QE_Opcode_PushFalse) {
      depth++;
    } else if (op == QE_Opcode_Halt) {
      break;
    } else {
      depth--;
    }
    if (depth > max_depth) max_depth = depth;
  }
  return (U32)max_depth;
}

B32 qe_bytecode_program_build(QE_Table* table, const QE_IRNode* nodes, U32 count, QE_Bytecode* out)
{
  memset(out, 0, sizeof(*out));
  if (!nodes || count == 0) {
    qe_emit(out, QE_Opcode_PushTrue);
  } else {
    if (qe_ir_subtree_end(nodes, 0) != count) {
      snprintf(out->error, sizeof(out->error), "malformed WHERE tree");
      return 0;
    }
    if (!qe_compile_condition(table, nodes, 0, out)) return 0;
  }
  qe_emit(out, QE_Opcode_Halt);
  if (out->overflow) {
    snprintf(out->error, sizeof(out->error), "WHERE too large for the bytecode buffers");
    return 0;
  }
  return 1;
}

/* ---------------- numeric AND leaves ---------------- */
static U32 qe_collect_and_nodes(const QE_IRNode* nodes, U32 i, U32* out, U32 n, U32 cap)
{
  if (nodes[i].type == QE_IR_OPERATOR && nodes[i].op == QE_IROP_AND && nodes[i].child_count == 2) {
    U32 left = i + 1, right = qe_ir_subtree_end(nodes, left);
    n = qe_collect_and_nodes(nodes, left, out, n, cap);
    return qe_collect_and_nodes(nodes, right, out, n, cap);
  }
  if (n < cap) out[n] = i;
  return n + 1;
}

/* qe_resolve_leaf_comparison: `column op literal` or `literal op column` */
static B32 qe_resolve_leaf(const QE_IRNode* nodes, U32 i, QE_Leaf* leaf)
{
  const QE_IRNode* n = &nodes[i];
  leaf->negate = 0;
  leaf->set = 0;
  /* `column IS [NOT] NULL` (query_exec.c:13902-13917) */
  if (n->type == QE_IR_OPERATOR && n->child_count == 1 && (n->op == QE_IROP_IS_NULL || n->op == QE_IROP_IS_NOT_NULL) &&
      nodes[i + 1].type == QE_IR_COLUMN) {
    leaf->column = nodes[i + 1].column;
    leaf->op = n->op == QE_IROP_IS_NULL ? QE_LEAF_IS_NULL : QE_LEAF_IS_NOT_NULL;
    leaf->target = 0.0;
    return 1;
  }
  if (n->type != QE_IR_OPERATOR || n->child_count != 2) return 0;
  U32 left = i + 1, right = qe_ir_subtree_end(nodes, left);
  const QE_IRNode *l = &nodes[left], *r = &nodes[right];
  U32 op;
  switch (n->op) {
    case QE_IROP_EQ: op = 0; break;
    case QE_IROP_LT: op = 1; break;
    case QE_IROP_LE: op = 2; break;
    case QE_IROP_GT: op = 3; break;
    case QE_IROP_GE: op = 4; break;
    default: return 0;
  }
  if (l->type == QE_IR_COLUMN && r->type == QE_IR_NUMERIC) {
    leaf->column = l->column;
    leaf->op = op;
    leaf->target = r->value;
    return 1;
  }
  if (l->type == QE_IR_NUMERIC && r->type == QE_IR_COLUMN) {
    static const U32 flip[5] = {0, 3, 4, 1, 2};
    leaf->column = r->column;
    leaf->op = flip[op];
    leaf->target = l->value;
    return 1;
  }
  return 0;
}

B32 qe_collect_and_leaves(const QE_IRNode* nodes, U32 count, QE_Leaf* leaves, U32* leaf_count, U32 max_leaves)
{
  if (!nodes || count == 0) return 0;
  U32 idx[QE_SCAN_MAX_LEAVES + 1];
  U32 n = qe_collect_and_nodes(nodes, 0, idx, 0, QE_SCAN_MAX_LEAVES + 1);
  if (n == 0 || n > max_leaves || n > QE_SCAN_MAX_LEAVES) return 0;
  for (U32 k = 0; k < n; k++)
    if (!qe_resolve_leaf(nodes, idx[k], &leaves[k])) return 0;
  *leaf_count = n;
  return 1;
}

/* ---------------- any AND / OR WHERE: disjunction of conjunctions ---------------- */
typedef struct QE_TermSet {
  U32 n;
  U8 t[QE_WHERE_MAX_TERMS];
} QE_TermSet;

static B32 qe_leaf_same(const QE_Leaf* a, const QE_Leaf* b)
{
  return a->column == b->column && a->op == b->op && a->negate == b->negate && a->set == b->set &&
         memcmp(&a->target, &b->target, sizeof(F64)) == 0;
}

/* index of the leaf in w (added when new), -1 when there is no room */
static S32 qe_where_leaf(QE_Where* w, const QE_Leaf* leaf)
{
  for (U32 i = 0; i < w->leaf_count; i++)
    if (qe_leaf_same(&w->leaf[i], leaf)) return (S32)i;
  if (w->leaf_count >= QE_SCAN_MAX_LEAVES) return -1;
  w->leaf[w->leaf_count] = *leaf;
  return (S32)w->leaf_count++;
}

static B32 qe_terms_add(QE_TermSet* s, U8 term)
{
  for (U32 i = 0; i < s->n; i++)
    if (s->t[i] == term) return 1;
  if (s->n >= QE_WHERE_MAX_TERMS) return 0;
  s->t[s->n++] = term;
  return 1;
}

static B32 qe_terms_leaf(QE_Where* w, const QE_Leaf* leaf, QE_TermSet* out)
{
  const S32 i = qe_where_leaf(w, leaf);
  out->n = 0;
  return i >= 0 && qe_terms_add(out, (U8)(1u << i));
}

static B32 qe_item_integral(F64 v, F64 lo, F64 hi) { return v == floor(v) && v >= lo && v <= hi; }

/* IN / NOT IN over a column: one set leaf when the column is a 32-bit
 * integer and the items that can match lie within 64 consecutive values;
 * otherwise an OR of = leaves (an AND of != leaves) */
static B32 qe_terms_in(const QE_Table* t, const QE_IRNode* nodes, U32 i, QE_Where* w, QE_TermSet* out)
{
  const QE_IRNode* n = &nodes[i];
  const U32 left = i + 1, right = qe_ir_subtree_end(nodes, left);
  if (n->child_count != 2 || nodes[left].type != QE_IR_COLUMN || nodes[right].type != QE_IR_INLIST) return 0;
  const U32 col = nodes[left].column;
  if (col >= t->column_count) return 0;
  const B32 is_not = n->op == QE_IROP_NOT_IN;
  const U32 items = nodes[right].child_count;
  for (U32 k = 0; k < items; k++)
    if (nodes[right + 1 + k].type != QE_IR_NUMERIC) return 0;
  QE_Leaf leaf;
  memset(&leaf, 0, sizeof(leaf));
  leaf.column = col;
  const U32 type = t->columns[col].type;
  if (type == CQL_ColumnType_U32 || type == CQL_ColumnType_I32) {
    const F64 vlo = type == CQL_ColumnType_U32 ? 0.0 : -2147483648.0;
    const F64 vhi = type == CQL_ColumnType_U32 ? 4294967295.0 : 2147483647.0;
    F64 mn = 0.0, mx = 0.0;
    U32 hits = 0;
    for (U32 k = 0; k < items; k++) {
      const F64 v = nodes[right + 1 + k].value;
      if (!qe_item_integral(v, vlo, vhi)) continue; /* matches no value of the column */
      mn = hits ? fmin(mn, v) : v;
      mx = hits ? fmax(mx, v) : v;
      hits++;
    }
    if (hits == 0 || mx - mn < 64.0) {
      /* the items no value equals drop out: IN keeps none, NOT IN passes
       * every non-NULL row */
      leaf.op = QE_LEAF_IN_SET;
      leaf.target = hits ? mn : 0.0;
      for (U32 k = 0; k < items; k++) {
        const F64 v = nodes[right + 1 + k].value;
        if (qe_item_integral(v, vlo, vhi)) leaf.set |= 1ull << (U32)(v - mn);
      }
      leaf.negate = is_not;
      return qe_terms_leaf(w, &leaf, out);
    }
  }
  out->n = 0;
  if (items == 0) {
    if (!is_not) return 1; /* IN (): never true */
    leaf.op = QE_LEAF_IS_NOT_NULL;
    return qe_terms_leaf(w, &leaf, out);
  }
  U8 all = 0;
  for (U32 k = 0; k < items; k++) {
    leaf.op = 0;
    leaf.target = nodes[right + 1 + k].value;
    leaf.negate = is_not;
    const S32 li = qe_where_leaf(w, &leaf);
    if (li < 0) return 0;
    if (is_not) all |= (U8)(1u << li);
    else if (!qe_terms_add(out, (U8)(1u << li))) return 0;
  }
  if (is_not) return qe_terms_add(out, all);
  return 1;
}

static B32 qe_terms_of(const QE_Table* t, const QE_IRNode* nodes, U32 i, QE_Where* w, QE_TermSet* out)
{
  const QE_IRNode* n = &nodes[i];
  out->n = 0;
  if (n->type == QE_IR_OPERATOR && (n->op == QE_IROP_AND || n->op == QE_IROP_OR) && n->child_count == 2) {
    const U32 left = i + 1, right = qe_ir_subtree_end(nodes, left);
    QE_TermSet a, b;
    if (!qe_terms_of(t, nodes, left, w, &a) || !qe_terms_of(t, nodes, right, w, &b)) return 0;
    if (n->op == QE_IROP_OR) {
      for (U32 k = 0; k < a.n; k++)
        if (!qe_terms_add(out, a.t[k])) return 0;
      for (U32 k = 0; k < b.n; k++)
        if (!qe_terms_add(out, b.t[k])) return 0;
      return 1;
    }
    for (U32 x = 0; x < a.n; x++)
      for (U32 y = 0; y < b.n; y++)
        if (!qe_terms_add(out, (U8)(a.t[x] | b.t[y]))) return 0;
    return 1;
  }
  if (n->type == QE_IR_OPERATOR && (n->op == QE_IROP_IN || n->op == QE_IROP_NOT_IN)) return qe_terms_in(t, nodes, i, w, out);
  QE_Leaf leaf;
  memset(&leaf, 0, sizeof(leaf));
  if (n->type == QE_IR_OPERATOR && n->op == QE_IROP_NE && n->child_count == 2) {
    /* a != b: the complement of a = b, NULL failing both */
    const U32 left = i + 1, right = qe_ir_subtree_end(nodes, left);
    const QE_IRNode *l = &nodes[left], *r = &nodes[right];
    if (l->type == QE_IR_COLUMN && r->type == QE_IR_NUMERIC) {
      leaf.column = l->column;
      leaf.target = r->value;
    } else if (l->type == QE_IR_NUMERIC && r->type == QE_IR_COLUMN) {
      leaf.column = r->column;
      leaf.target = l->value;
    } else {
      return 0;
    }
    leaf.op = 0;
    leaf.negate = 1;
  } else if (!qe_resolve_leaf(nodes, i, &leaf)) {
    return 0;
  }
  if (leaf.column >= t->column_count) return 0;
  return qe_terms_leaf(w, &leaf, out);
}

B32 qe_where_build(const QE_Table* t, const QE_IRNode* nodes, U32 count, QE_Where* out)
{
  memset(out, 0, sizeof(*out));
  if (!nodes || count == 0) return 1;
  if (qe_ir_subtree_end(nodes, 0) != count) return 0;
  QE_TermSet s;
  if (!qe_terms_of(t, nodes, 0, out, &s)) return 0;
  if (s.n == 0) {
    /* never true: one leaf no row passes (= NaN) */
    QE_Leaf never;
    memset(&never, 0, sizeof(never));
    never.column = ~0u;
    for (U32 k = 0; k < count && never.column == ~0u; k++)
      if (nodes[k].type == QE_IR_COLUMN) never.column = nodes[k].column;
    never.target = NAN;
    out->leaf_count = 1;
    out->leaf[0] = never;
    out->term_count = 0;
    return never.column < t->column_count;
  }
  /* absorption: a term holding every leaf of another adds no row */
  U8 kept[QE_WHERE_MAX_TERMS];
  U32 nk = 0;
  for (U32 a = 0; a < s.n; a++) {
    B32 redundant = 0;
    for (U32 b = 0; b < s.n && !redundant; b++)
      if (b != a && (s.t[a] & s.t[b]) == s.t[b] && s.t[a] != s.t[b]) redundant = 1;
    if (!redundant) kept[nk++] = s.t[a];
  }
  /* leaves no kept term reads are dropped, and the rest renumbered */
  U8 used = 0;
  for (U32 k = 0; k < nk; k++) used |= kept[k];
  U32 map[QE_SCAN_MAX_LEAVES];
  U32 m = 0;
  for (U32 l = 0; l < out->leaf_count; l++) {
    map[l] = m;
    if (used & (1u << l)) out->leaf[m++] = out->leaf[l];
  }
  out->leaf_count = m;
  for (U32 k = 0; k < nk; k++) {
    U8 r = 0;
    for (U32 l = 0; l < QE_SCAN_MAX_LEAVES; l++)
      if (kept[k] & (1u << l)) r |= (U8)(1u << map[l]);
    kept[k] = r;
  }
  const U8 all = (U8)((1u << m) - 1u);
  if (nk == 1 && kept[0] == all) {
    out->term_count = 0;
  } else {
    out->term_count = nk;
    memcpy(out->term, kept, nk);
  }
  return 1;
}

U32 qe_where_words(const QE_Table* t, const QE_Where* w, const U32* slot, B32 nulls, B32 enc, U32* words, U32* packed)
{
  memset(words, 0, QE_WHERE_WORDS * 4);
  for (U32 l = 0; l < w->leaf_count; l++) {
    const QE_Leaf* f = &w->leaf[l];
    U32* lw = &words[6 * l];
    if (f->op == QE_LEAF_IN_SET) {
      /* the set's lowest value as an integer interval's lo (the i32 bias
       * and encodings apply to it as to an interval); the bits in words
       * 2 and 4 */
      const U32 type = t->columns[f->column].type;
      const U32 bias = type == CQL_ColumnType_I32 ? QE_SCAN_INT_BIAS : 0u;
      lw[0] = slot[l];
      lw[1] = type | QE_SCAN_INT_FLAG | QE_SCAN_INT_SET | bias | (f->set ? 0u : QE_SCAN_INT_NEVER);
      lw[2] = (U32)f->set;
      lw[3] = type == CQL_ColumnType_I32 ? ((U32)(S32)f->target ^ 0x80000000u) : (U32)f->target;
      lw[4] = (U32)(f->set >> 32);
      lw[5] = 0;
    } else {
      qe_simple_leaf_words(lw, slot[l], t->columns[f->column].type, f->op, f->target);
    }
    if (f->negate) lw[1] |= QE_SCAN_NEGATE;
    if (nulls) qe_leaf_nulls(lw, t, f->column);
    if (enc) qe_leaf_encode(lw, qe_col_enc(t, f->column), t->columns[f->column].enc_base);
  }
  for (U32 k = 0; k < w->term_count; k++) words[6 * QE_SCAN_MAX_LEAVES + k / 4] |= (U32)w->term[k] << (8 * (k % 4));
  *packed = w->leaf_count | (w->term_count << 8);
  return w->term_count ? QE_WHERE_WORDS * 4 : 24 * w->leaf_count;
}

U32 qe_collect_prune_leaves(const QE_IRNode* nodes, U32 count, QE_Leaf* leaves, U32 max_leaves)
{
  if (!nodes || count == 0) return 0;
  U32 idx[64];
  U32 n = qe_collect_and_nodes(nodes, 0, idx, 0, 64);
  if (n > 64) n = 64;
  U32 out = 0;
  for (U32 k = 0; k < n && out < max_leaves; k++)
    if (qe_resolve_leaf(nodes, idx[k], &leaves[out])) out++;
  return out;
}

/* query_exec.c:1378-1404 */
B32 qe_scan_int_interval(U32 op, F64 target, F64 code_offset, F64 max_code, U32* out_lo, U32* out_span)
{
  if (target != target) return 0;
  F64 low = -1e18, high = 1e18;
  switch (op) {
    case 0:
      if (floor(target) != target) return 0;
      low = target;
      high = target;
      break;
    case 1: high = ceil(target) - 1.0; break;
    case 2: high = floor(target); break;
    case 3: low = floor(target) + 1.0; break;
    case 4: low = ceil(target); break;
    default: return 0;
  }
  low = fmax(low - code_offset, 0.0);
  high = fmin(high - code_offset, max_code);
  if (low > high) return 0;
  *out_lo = (U32)low;
  *out_span = (U32)(high - low);
  return 1;
}

/* query_exec.c:1408-1456, full-width columns (width 0; qe_leaf_encode
 * rewrites a leaf for an encoded column) */
void qe_simple_leaf_words(U32* words, U32 slot, U32 type, U32 op, F64 target)
{
  U64 bits;
  memcpy(&bits, &target, 8);
  words[0] = slot;
  words[1] = type;
  words[2] = op;
  words[3] = (U32)(bits & 0xffffffffu);
  words[4] = (U32)(bits >> 32);
  words[5] = 0;
  if (op >= QE_LEAF_IS_NULL) {
    words[1] = type | QE_SCAN_NULL_TEST;
    words[3] = words[4] = 0;
    return;
  }
  if ((type == CQL_ColumnType_U32 || type == CQL_ColumnType_I32) && qe_settings_u64("QE_SCAN_INT_COMPARE", 1) != 0) {
    F64 code_offset = 0.0, max_code = 4294967295.0;
    U32 flags = QE_SCAN_INT_FLAG;
    if (type == CQL_ColumnType_I32) {
      code_offset = -2147483648.0;
      flags |= QE_SCAN_INT_BIAS;
    }
    U32 lo = 0, span = 0;
    if (!qe_scan_int_interval(op, target, code_offset, max_code, &lo, &span)) flags |= QE_SCAN_INT_NEVER;
    words[1] = type | flags;
    words[3] = lo;
    words[4] = span;
    words[5] = 0;
  }
}

/* ---------------- zone maps ---------------- */
static B32 qe_zone_prunable(const QE_Leaf* leaf, F64 zmin, F64 zmax)
{
  if (leaf->op >= QE_LEAF_IS_NULL) return 0; /* zone maps hold no NULL counts */
  switch (leaf->op) {
    case 0: return leaf->target < zmin || leaf->target > zmax;
    case 1: return zmin >= leaf->target;
    case 2: return zmin > leaf->target;
    case 3: return zmax <= leaf->target;
    default: return zmax < leaf->target;
  }
}

void qe_zonemap_build_super(QE_Column* col)
{
  free(col->super_zone);
  col->super_zone = 0;
  col->super_count = 0;
  if (!col->zone_count || !col->zone_min || !col->zone_max) return;
  U64 n = (col->zone_count + (1ull << QE_SUPER_SHIFT) - 1) >> QE_SUPER_SHIFT;
  F64* sz = (F64*)malloc(sizeof(F64) * 4 * n);
  if (!sz) return;
  for (U64 g = 0; g < n; g++) {
    U64 z0 = g << QE_SUPER_SHIFT, z1 = z0 + (1ull << QE_SUPER_SHIFT);
    if (z1 > col->zone_count) z1 = col->zone_count;
    F64 mnmn = col->zone_min[z0], mxmn = col->zone_min[z0], mnmx = col->zone_max[z0], mxmx = col->zone_max[z0];
    for (U64 z = z0 + 1; z < z1; z++) {
      F64 a = col->zone_min[z], b = col->zone_max[z];
      if (a < mnmn) mnmn = a;
      if (a > mxmn) mxmn = a;
      if (b < mnmx) mnmx = b;
      if (b > mxmx) mxmx = b;
    }
    sz[4 * g] = mnmn;
    sz[4 * g + 1] = mxmn;
    sz[4 * g + 2] = mnmx;
    sz[4 * g + 3] = mxmx;
  }
  col->super_zone = sz;
  col->super_count = n;
}

/* a group of zones against one leaf: 1 every zone prunable, 0 none
 * prunable, -1 undecided (each outcome implies the zone-by-zone one) */
static int qe_super_state(const QE_Leaf* leaf, const F64* g)
{
  const F64 t = leaf->target, mnmn = g[0], mxmn = g[1], mnmx = g[2], mxmx = g[3];
  if (leaf->op >= QE_LEAF_IS_NULL) return 0;
  switch (leaf->op) {
    case 0:
      if (t < mnmn || t > mxmx) return 1;
      return (mxmn <= t && t <= mnmx) ? 0 : -1;
    case 1:
      if (mnmn >= t) return 1;
      return mxmn < t ? 0 : -1;
    case 2:
      if (mnmn > t) return 1;
      return mxmn <= t ? 0 : -1;
    case 3:
      if (mxmx <= t) return 1;
      return mnmx > t ? 0 : -1;
    default:
      if (mxmx < t) return 1;
      return mnmx >= t ? 0 : -1;
  }
}

void qe_zonemap_prune(QE_Table* table, const QE_Leaf* leaves, U32 leaf_count, QE_Prune* out)
{
  memset(out, 0, sizeof(*out));
  out->row_begin = 0;
  out->row_end = table->row_count;
  U64 zones = (table->row_count + (1ull << QE_ZONE_SHIFT) - 1) >> QE_ZONE_SHIFT;
  out->zones_total = zones;
  out->zones_left = zones;
  if (leaf_count == 0 || zones == 0 || qe_settings_u64("QE_ZONEMAP", 1) == 0) return;
  for (U32 l = 0; l < leaf_count; l++)
    if (leaves[l].column >= table->column_count || !table->columns[leaves[l].column].zonemap_valid) return;
  U64 words = (zones + 31) / 32;
  U32* bits = (U32*)calloc(words, sizeof(U32));
  if (!bits) return;
  U64 left = 0, first = zones, last = 0;
  /* QE_ZONEMAP_SUPER (1): zones in groups of 64, a group decided by its
   * group values when they settle every zone of it */
  B32 super = qe_settings_u64("QE_ZONEMAP_SUPER", 1) != 0;
  for (U32 l = 0; l < leaf_count && super; l++) {
    const QE_Column* c = &table->columns[leaves[l].column];
    super = c->super_zone && c->super_count == ((zones + (1ull << QE_SUPER_SHIFT) - 1) >> QE_SUPER_SHIFT);
  }
  const U64 gz = 1ull << QE_SUPER_SHIFT;
  for (U64 z0 = 0; z0 < zones; z0 += gz) {
    U64 z1 = z0 + gz < zones ? z0 + gz : zones;
    int state = super ? 0 : -1;
    for (U32 l = 0; l < leaf_count && super; l++) {
      int s = qe_super_state(&leaves[l], table->columns[leaves[l].column].super_zone + 4 * (z0 >> QE_SUPER_SHIFT));
      if (s == 1) {
        state = 1;
        break;
      }
      if (s < 0) state = -1;
    }
    if (state == 1) {
      for (U64 z = z0; z < z1; z++) bits[z >> 5] |= 1u << (z & 31);
      continue;
    }
    if (state == 0) {
      left += z1 - z0;
      if (z0 < first) first = z0;
      last = z1 - 1;
      continue;
    }
    for (U64 z = z0; z < z1; z++) {
      B32 skip = 0;
      for (U32 l = 0; l < leaf_count && !skip; l++) {
        QE_Column* c = &table->columns[leaves[l].column];
        skip = qe_zone_prunable(&leaves[l], c->zone_min[z], c->zone_max[z]);
      }
      if (skip) bits[z >> 5] |= 1u << (z & 31);
      else {
        left++;
        if (z < first) first = z;
        last = z;
      }
    }
  }
  out->zones_left = left;
  if (left == zones) {
    free(bits);
    return;
  }
  out->skip_bits = bits;
  out->skip_words = words;
  if (left == 0) {
    out->row_begin = out->row_end = 0;
    return;
  }
  out->row_begin = first << QE_ZONE_SHIFT;
  U64 end = (last + 1) << QE_ZONE_SHIFT;
  out->row_end = end < table->row_count ? end : table->row_count;
}

void qe_prune_free(QE_Prune* p)
{
  free(p->skip_bits);
  p->skip_bits = 0;
}

/* ---------------- the scan driver ---------------- */
/* the slot columns; encoded copies where `encoded` (kernels whose leaves
 * were rewritten for codes) */
static void qe_bind_columns(GPU_Kernel* k, QE_Table* t, const QE_ColumnSlots* slots, B32 encoded)
{
  for (U32 s = 0; s < QE_MAX_COLUMN_BINDINGS; s++) {
    if (s < slots->count)
      gpu_kernel_set_arg_buffer(k, 6 + s, encoded ? qe_col_kbuf(t, slots->column[s]) : t->columns[slots->column[s]].buffer);
    else gpu_hip_kernel_clear_arg(k, 6 + s);
  }
}

static B32 qe_scan_impl(QE_Table* table, const QE_IRNode* where, U32 node_count, B32 count_only, U64* out_ids,
                        U64 cap, U64* out_count, QE_Trace* trace, GPU_Buffer** device_ids);

B32 qe_scan_filter(QE_Table* table, const QE_IRNode* where, U32 node_count, B32 count_only, U64* out_ids, U64 cap,
                   U64* out_count, QE_Trace* trace)
{
  if (!qe_where_non_null(table, where, node_count, "qe_scan_filter")) return 0;
  /* host-backed rows stream through the chunk pipeline; with QE_SCAN_STREAM
   * 0 (query_exec.c:2798) the kernels read them in place over the host link */
  if (table && out_count && table->host_backed) {
    if (qe_settings_u64("QE_SCAN_STREAM", 1) != 0)
      return qe_stream_scan(table, where, node_count, count_only, out_ids, cap, out_count, trace);
    if (!qe_require_device(table, "qe_scan_filter")) return 0;
  }
  return qe_scan_impl(table, where, node_count, count_only, out_ids, cap, out_count, trace, 0);
}

/* the matching row ids left on the device (a QE_DeviceSelection,
 * query_exec.h:387-400): *out_ids is a scratch buffer of *out_count U64 ids
 * the caller releases with qe_scratch_release (0 when nothing matched) */
B32 qe_scan_filter_device(QE_Table* table, const QE_IRNode* where, U32 node_count, GPU_Buffer** out_ids, U64* out_count)
{
  *out_ids = 0;
  if (!qe_require_device(table, "qe_scan_filter_device") || !qe_where_non_null(table, where, node_count, "qe_scan_filter"))
    return 0;
  return qe_scan_impl(table, where, node_count, 0, 0, ~0ull, out_count, 0, out_ids);
}

static B32 qe_scan_impl(QE_Table* table, const QE_IRNode* where, U32 node_count, B32 count_only, U64* out_ids,
                        U64 cap, U64* out_count, QE_Trace* trace, GPU_Buffer** device_ids)
{
  U64 t0, p0;
  qe_trace_begin(trace, &t0, &p0);
  if (!table || !out_count) {
    qe_set_error("qe_scan_filter: null argument");
    return 0;
  }
  U64 rows = table->row_count;
  if (trace) trace->rows_before = rows;
  static __thread QE_Bytecode prog;
  if (!qe_bytecode_program_build(table, where, node_count, &prog)) {
    qe_set_error("qe_scan_filter: %s", prog.error);
    return 0;
  }
  /* constant predicates (no column read); IS [NOT] NULL over a column
   * with NULLs is not constant */
  B32 reads_nulls = 0;
  for (U32 i = 0; where && i < node_count; i++)
    if (where[i].type == QE_IR_COLUMN && where[i].column < table->column_count && table->columns[where[i].column].has_nulls)
      reads_nulls = 1;
  if (!reads_nulls && prog.word_count == 2 && (prog.words[0] == QE_Opcode_PushTrue || prog.words[0] == QE_Opcode_PushFalse)) {
    U64 n = prog.words[0] == QE_Opcode_PushTrue ? rows : 0;
    *out_count = n;
    if (!count_only && out_ids)
      for (U64 i = 0; i < n && i < cap; i++) out_ids[i] = i;
    if (device_ids && n) {
      U64* tmp = (U64*)malloc(sizeof(U64) * n);
      GPU_Buffer* ids = tmp ? qe_scratch("scan_ids_dev", n * 8) : 0;
      if (!ids) {
        free(tmp);
        qe_set_error("qe_scan_filter: out of memory");
        return 0;
      }
      for (U64 i = 0; i < n; i++) tmp[i] = i;
      const B32 wrote = qe_buffer_write_checked(ids, tmp, n * 8);
      free(tmp);
      if (!wrote) {
        qe_scratch_release(ids);
        return 0;
      }
      *device_ids = ids;
    }
    if (trace) {
      trace->path = QE_ScanPath_Constant;
      trace->rows_after = n;
    }
    qe_trace_end(trace, t0, p0);
    return 1;
  }

  QE_Leaf leaves[QE_SCAN_MAX_LEAVES];
  U32 leaf_count = 0;
  B32 is_and = qe_collect_and_leaves(where, node_count, leaves, &leaf_count, QE_SCAN_MAX_LEAVES);
  /* the library scan_simple evaluates any AND / OR of at most 8 leaves */
  static __thread QE_Where wh;
  const B32 leaf_form = qe_where_build(table, where, node_count, &wh);
  QE_Leaf prune_leaves[32];
  U32 prune_count = qe_collect_prune_leaves(where, node_count, prune_leaves, 32);
  QE_Prune prune;
  qe_zonemap_prune(table, prune_leaves, prune_count, &prune);
  if (trace) trace->rows_scanned = prune.row_end - prune.row_begin;
  if (prune.zones_left == 0) {
    qe_prune_free(&prune);
    *out_count = 0;
    if (trace) trace->path = QE_ScanPath_ZoneEmpty;
    qe_trace_end(trace, t0, p0);
    return 1;
  }

  /* kernel choice, in the order of query_exec.c:2566-2656 */
  U32 path;
  B32 force_interp = qe_settings_u64("QE_SCAN_FORCE_INTERPRETER", 0) != 0;
  B32 perquery = qe_settings_u64("QE_PERQUERY_KERNELS", 1) != 0 && rows >= qe_settings_u64("QE_PERQUERY_MIN_ROWS", 0);
  B32 simple_ok = is_and && qe_settings_u64("QE_SCAN_SIMPLE_KERNEL", 1) != 0;
  /* NULLs are evaluated by the library scan_simple only */
  B32 nullable = 0;
  for (U32 l = 0; l < wh.leaf_count && leaf_form; l++)
    nullable |= table->columns[wh.leaf[l].column].has_nulls || wh.leaf[l].op == QE_LEAF_IS_NULL ||
                wh.leaf[l].op == QE_LEAF_IS_NOT_NULL;
  if (nullable) path = QE_ScanPath_Simple;
  else if (force_interp) path = QE_ScanPath_Interpreter;
  else if (simple_ok) path = perquery && qe_settings_u64("QE_PERQUERY_SIMPLE", 0) ? QE_ScanPath_GeneratedSimple : QE_ScanPath_Simple;
  else if (perquery) path = QE_ScanPath_Generated;
  else path = QE_ScanPath_Interpreter;
  /* the interpreter's stack holds QE_SCAN_MAX_STACK (8) values
   * (query_exec.h:41); deeper programs need a generated kernel */
  if (path == QE_ScanPath_Interpreter && qe_bytecode_max_stack_depth(&prog) > 8) {
    if (perquery && !force_interp) path = QE_ScanPath_Generated;
    else {
      qe_prune_free(&prune);
      qe_set_error("qe_scan_filter: WHERE needs a stack deeper than 8 and no generated kernel is allowed");
      return 0;
    }
  }
  /* QE_ROUTE_CONJUNCTIONS: a new shape's first run on the interpreter */
  if (path == QE_ScanPath_Generated && qe_bytecode_max_stack_depth(&prog) <= 8 &&
      qe_settings_u64("QE_SCAN_INTERPRETER", 1) != 0 && qk_route_defer(&prog))
    path = QE_ScanPath_Interpreter;
  if (path == QE_ScanPath_Interpreter && qe_settings_u64("QE_SCAN_INTERPRETER", 1) == 0) {
    qe_prune_free(&prune);
    qe_set_error("qe_scan_filter: no GPU kernel for this WHERE (interpreter disabled)");
    return 0;
  }

  GPU_Kernel* k = 0;
  QE_ColumnSlots slots;
  memset(&slots, 0, sizeof(slots));
  U32 leaf_words[QE_WHERE_WORDS];
  U32 leaf_packed = leaf_count, leaf_bytes = 24 * leaf_count;
  F64 consts[256];
  U32 const_count = 0;
  U64 compile_ns = 0;
  if (path == QE_ScanPath_GeneratedSimple) {
    k = qk_simple_kernel(table, leaves, leaf_count, &slots, &compile_ns);
    if (!k) path = QE_ScanPath_Simple; /* the library kernel until a compile lands */
  }
  if (path == QE_ScanPath_Generated) {
    k = qk_scan_kernel(table, where, node_count, &slots, consts, &const_count, &compile_ns);
    if (!k) {
      if (qe_bytecode_max_stack_depth(&prog) > 8 || qe_settings_u64("QE_SCAN_INTERPRETER", 1) == 0) {
        qe_prune_free(&prune);
        return 0; /* the generator's error stays set */
      }
      path = QE_ScanPath_Interpreter;
    }
  }
  if (path == QE_ScanPath_Simple || path == QE_ScanPath_GeneratedSimple) {
    if (path == QE_ScanPath_Simple) {
      memset(&slots, 0, sizeof(slots));
      k = qe_kernel(leaf_form && wh.term_count ? "scan_simple_or" : "scan_simple");
    }
    if (path == QE_ScanPath_Simple) {
      /* scan_simple reads encoded columns (a3) as codes */
      U32 slot[QE_SCAN_MAX_LEAVES];
      for (U32 l = 0; l < wh.leaf_count; l++) slot[l] = qe_slot_of(&slots, wh.leaf[l].column);
      leaf_bytes = qe_where_words(table, &wh, slot, 1, 1, leaf_words, &leaf_packed);
    }
    /* the generated form reads full-width columns without NULLs */
    for (U32 l = 0; l < leaf_count && path == QE_ScanPath_GeneratedSimple; l++) {
      U32 slot = qe_slot_of(&slots, leaves[l].column);
      qe_simple_leaf_words(&leaf_words[6 * l], slot, table->columns[leaves[l].column].type, leaves[l].op, leaves[l].target);
    }
  }
  if (path == QE_ScanPath_Interpreter) {
    k = qe_kernel("scan_filter");
    slots = prog.slots;
  }
  if (!k) {
    qe_prune_free(&prune);
    return 0;
  }

  /* device buffers: program, constants, skip bits, counter, block counts;
   * the uploads ride the first batch (pinned staging, no host sync) */
  GPU_Buffer* prog_buf = 0;
  GPU_Buffer* const_buf = 0;
  const void* prog_src = 0;
  U64 prog_bytes = 0;
  const void* const_src = 0;
  U64 const_bytes = 0;
  if (path == QE_ScanPath_Simple || path == QE_ScanPath_GeneratedSimple) {
    prog_buf = qe_scratch("scan_leaves", sizeof(leaf_words));
    prog_src = leaf_words;
    prog_bytes = leaf_bytes;
  } else if (path == QE_ScanPath_Interpreter) {
    prog_buf = qe_scratch("scan_bytecode", prog.word_count * 4);
    const_buf = qe_scratch("scan_consts", (prog.const_count ? prog.const_count : 1) * 8);
    prog_src = prog.words;
    prog_bytes = prog.word_count * 4;
    const_src = prog.consts;
    const_bytes = prog.const_count * 8;
  } else {
    const_buf = qe_scratch("scan_consts", (const_count ? const_count : 1) * 8);
    const_src = consts;
    const_bytes = const_count * 8;
  }
  GPU_Buffer* skip_buf = prune.skip_bits ? qe_scratch("scan_skip", prune.skip_words * 4) : 0;
  U64 scan_rows = prune.row_end - prune.row_begin;
  U64 quads = (prune.row_end + 3) / 4 - prune.row_begin / 4;
  U32 blocks = qe_stream_blocks(quads);
  GPU_Buffer* counter = qe_scratch("scan_counter", 64);
  GPU_Buffer* block_counts = qe_scratch("scan_block_counts", (U64)blocks * 8);
  B32 ok = counter && block_counts && (prog_buf || path == QE_ScanPath_Generated) && (!prune.skip_bits || skip_buf) &&
           (!const_src || const_buf);
  if (ok) {
    if (prog_buf) gpu_kernel_set_arg_buffer(k, 0, prog_buf);
    else gpu_hip_kernel_clear_arg(k, 0);
    if (const_buf) gpu_kernel_set_arg_buffer(k, 1, const_buf);
    else gpu_hip_kernel_clear_arg(k, 1);
    gpu_kernel_set_arg_buffer(k, 2, block_counts);
    gpu_kernel_set_arg_buffer(k, 4, counter);
    if (skip_buf) gpu_kernel_set_arg_buffer(k, 5, skip_buf);
    else gpu_hip_kernel_clear_arg(k, 5);
    qe_bind_columns(k, table, &slots, path == QE_ScanPath_Simple);
    gpu_kernel_set_arg_u64(k, 0, scan_rows);
    gpu_kernel_set_arg_u64(k, 3, QE_ZONE_SHIFT);
    gpu_kernel_set_arg_u64(k, 4, skip_buf ? 1 : 0);
    gpu_kernel_set_arg_u64(k, 5, prune.row_begin);
    gpu_kernel_set_arg_u64(k, 7, leaf_packed);
    gpu_hip_kernel_clear_arg(k, 3);
  }
  U64 total = 0;
  U32 kernels = 0;
  GPU_Batch* b = ok ? gpu_batch_begin(prog_bytes + const_bytes, (U64)blocks * 8) : 0;
  if (ok && b) {
    if (prog_bytes) gpu_batch_buffer_write(b, prog_buf, (void*)prog_src, prog_bytes);
    if (const_bytes) gpu_batch_buffer_write(b, const_buf, (void*)const_src, const_bytes);
    if (skip_buf) gpu_batch_buffer_write(b, skip_buf, prune.skip_bits, prune.skip_words * 4);
  }
  if (ok && count_only) {
    /* count mode: one count per block, added here (as the reference adds
     * its per-chunk counts, query_exec.c:2184-2190) */
    U64* counts = (U64*)malloc(sizeof(U64) * blocks);
    gpu_kernel_set_arg_u64(k, 6, 0); /* QK_MODE_COUNT */
    gpu_batch_kernel_execute(b, k, blocks * QE_WG, QE_WG);
    if (counts) gpu_batch_buffer_read(b, block_counts, counts, (U64)blocks * 8);
    ok = gpu_batch_end(b) && counts;
    if (ok)
      for (U32 i = 0; i < blocks; i++) total += counts[i];
    free(counts);
    kernels = 1;
  } else if (ok) {
    U64* counts = (U64*)malloc(sizeof(U64) * blocks);
    gpu_kernel_set_arg_u64(k, 6, 1); /* QK_MODE_BLOCK_COUNTS */
    gpu_batch_kernel_execute(b, k, blocks * QE_WG, QE_WG);
    if (counts) gpu_batch_buffer_read(b, block_counts, counts, (U64)blocks * 8);
    ok = gpu_batch_end(b) && counts;
    kernels = 1;
    if (ok) {
      for (U32 i = 0; i < blocks; i++) total += counts[i];
      U64 want = total < cap ? total : cap;
      if (want > 0 && (out_ids || device_ids)) {
        GPU_Buffer* ids = qe_scratch(device_ids ? "scan_ids_dev" : "scan_ids", want * 8);
        ok = ids != 0;
        if (ok) {
          gpu_kernel_set_arg_buffer(k, 3, ids);
          gpu_kernel_set_arg_u64(k, 6, 2); /* QK_MODE_WRITE_IDS */
          b = gpu_batch_begin(0, want * 8);
          /* QE_SCAN_IDS_SENTINEL: pre-fill the output so an id the write
           * pass misses shows up as all ones */
          if (qe_settings_u64("QE_SCAN_IDS_SENTINEL", 0)) gpu_batch_buffer_fill(b, ids, want * 8, 0xffffffffu);
          gpu_batch_kernel_execute(b, k, blocks * QE_WG, QE_WG);
          if (!device_ids) gpu_batch_buffer_read(b, ids, out_ids, want * 8);
          ok = gpu_batch_end(b);
          kernels = 2;
          gpu_hip_kernel_clear_arg(k, 3);
        }
        if (device_ids && ok) *device_ids = ids;
        else qe_scratch_release(ids);
      }
    }
    free(counts);
  } else if (b) {
    gpu_batch_end(b);
  }
  qe_scratch_release(prog_buf);
  qe_scratch_release(const_buf);
  qe_scratch_release(skip_buf);
  qe_scratch_release(counter);
  qe_scratch_release(block_counts);
  qe_prune_free(&prune);
  if (!ok) {
    qe_set_error("qe_scan_filter: GPU dispatch failed: %s", gpu_hip_last_error());
    return 0;
  }
  *out_count = total;
  if (trace) {
    trace->path = path;
    trace->rows_after = total;
    trace->kernel_count = kernels;
    trace->compile_ns = compile_ns;
    trace->bytes_read = scan_rows * 4 * (U64)slots.count; /* every slot column is read once */
    snprintf(trace->kernel, sizeof(trace->kernel), "%s",
             path == QE_ScanPath_Interpreter ? "scan_filter" : (path == QE_ScanPath_Simple ? "scan_simple" : "qk_main"));
  }
  qe_trace_end(trace, t0, p0);
  return 1;
}
