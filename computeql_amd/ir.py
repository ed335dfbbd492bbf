"""WHERE trees in the planner's form: IR nodes flattened in prefix order.

The C ABI takes `QE_IRNode[]` (type, op, column, child_count, value); this
module builds those arrays from Python expressions, e.g.

    w = (col(2) < 200) & ((col(1) == 7) | (col(3) >= 4e9))

The same array layout is what the CPU oracle evaluates, so one expression
drives both sides of a parity test.
"""
from __future__ import annotations

import numpy as np

IR_OPERATOR, IR_COLUMN, IR_NUMERIC, IR_INLIST = 1, 2, 3, 4
OP_AND, OP_OR, OP_EQ, OP_NE, OP_LT, OP_GT, OP_LE, OP_GE = 1, 2, 3, 4, 5, 6, 7, 8
OP_IS_NULL, OP_IS_NOT_NULL, OP_IN, OP_NOT_IN = 9, 10, 11, 12

IR_DTYPE = np.dtype([("type", "<u4"), ("op", "<u4"), ("column", "<u4"), ("child_count", "<u4"), ("value", "<f8")])
assert IR_DTYPE.itemsize == 24


class Expr:
    def _cmp(self, other, op):
        return Op(op, [self, _wrap(other)])

    def __lt__(self, o):
        return self._cmp(o, OP_LT)

    def __le__(self, o):
        return self._cmp(o, OP_LE)

    def __gt__(self, o):
        return self._cmp(o, OP_GT)

    def __ge__(self, o):
        return self._cmp(o, OP_GE)

    def __eq__(self, o):  # noqa: D105 - builds a node, like an SQL '='
        return self._cmp(o, OP_EQ)

    def __ne__(self, o):
        return self._cmp(o, OP_NE)

    def __and__(self, o):
        return Op(OP_AND, [self, _wrap(o)])

    def __or__(self, o):
        return Op(OP_OR, [self, _wrap(o)])

    __hash__ = object.__hash__

    def isin(self, values, negate=False):
        return Op(OP_NOT_IN if negate else OP_IN, [self, InList([float(v) for v in values])])

    def is_null(self):
        return Op(OP_IS_NULL, [self])

    def is_not_null(self):
        return Op(OP_IS_NOT_NULL, [self])

    def between(self, lo, hi):
        return (self >= lo) & (self <= hi)

    def nodes(self) -> np.ndarray:
        """The flattened tree, built on first use and kept: an expression is
        not changed after it is built (the operators return new nodes), so
        a query run many times converts its WHERE once.  Read-only."""
        cached = self.__dict__.get("_nodes")
        if cached is None:
            out: list[tuple] = []
            self._emit(out)
            cached = np.array(out, dtype=IR_DTYPE)
            cached.flags.writeable = False
            self._nodes = cached
        return cached


class Column(Expr):
    def __init__(self, index: int):
        self.index = int(index)

    def _emit(self, out):
        out.append((IR_COLUMN, 0, self.index, 0, 0.0))

    def __repr__(self):
        return f"col({self.index})"


class Literal(Expr):
    def __init__(self, value: float):
        self.value = float(value)

    def _emit(self, out):
        out.append((IR_NUMERIC, 0, 0, 0, self.value))

    def __repr__(self):
        return repr(self.value)


class InList(Expr):
    def __init__(self, values):
        self.values = list(values)

    def _emit(self, out):
        out.append((IR_INLIST, 0, 0, len(self.values), 0.0))
        for v in self.values:
            out.append((IR_NUMERIC, 0, 0, 0, float(v)))


class Op(Expr):
    def __init__(self, op: int, children):
        self.op = op
        self.children = list(children)

    def _emit(self, out):
        out.append((IR_OPERATOR, self.op, 0, len(self.children), 0.0))
        for c in self.children:
            c._emit(out)

    def __repr__(self):
        names = {OP_AND: "AND", OP_OR: "OR", OP_EQ: "=", OP_NE: "!=", OP_LT: "<", OP_GT: ">", OP_LE: "<=",
                 OP_GE: ">=", OP_IN: "IN", OP_NOT_IN: "NOT IN", OP_IS_NULL: "IS NULL", OP_IS_NOT_NULL: "IS NOT NULL"}
        return f"({f' {names.get(self.op, self.op)} '.join(map(repr, self.children))})"


def _wrap(v):
    return v if isinstance(v, Expr) else Literal(v)


def col(index: int) -> Column:
    return Column(index)


def lit(value: float) -> Literal:
    return Literal(value)


def to_nodes(where) -> np.ndarray | None:
    if where is None:
        return None
    if isinstance(where, np.ndarray):
        return where.astype(IR_DTYPE, copy=False)
    return where.nodes()
