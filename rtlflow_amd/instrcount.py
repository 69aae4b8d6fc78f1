"""Instruction-count cost estimate of a logic vertex.

Restates V3InstrCount (src/V3InstrCount.cpp:30-35, 121-252) with the per-node
weights of src/V3AstNodes.h / src/V3Ast.h:1493-1506:

* every node costs ``widthInstrs`` = words for wide values, else 1;
* a variable read costs widthInstrs * instrCountLd (2), a write widthInstrs;
* multiply x3, divide/modulo x10, reduction-xor 1 + log2(width);
* ``&&``/``||`` add a branch (4); if/else and ?: cost the condition, a branch
  (4) and the larger of the two arms only;
* a select or memory read does not count the (possibly huge) vector it selects
  from, only its index;
* concatenation is free (its parts' cost is the destination's width);
* extensions are free.

This is the fallback cost the learned cost model replaces
(src/V3Partition.cpp:456 LogicMTask cost, :2486 finalizeCosts).
"""
from __future__ import annotations

from .ir import Assign, Case, Concat, Cond, Const, Expr, If, MemRd, Op, Ref, Sel, Stmt

BRANCH = 4
LOAD = 2
MUL = 3
DIV = 10


def _wi(width: int) -> int:
    return (width + 31) // 32 if width > 64 else 1


def expr_cost(e: Expr) -> int:
    if isinstance(e, Const):
        return _wi(e.width)
    if isinstance(e, Ref):
        return _wi(e.width) * LOAD
    if isinstance(e, MemRd):
        return _wi(e.width) + expr_cost(e.idx)
    if isinstance(e, Sel):
        return _wi(e.width) + (0 if isinstance(e.lsb, int) else expr_cost(e.lsb))
    if isinstance(e, Concat):
        return 0
    if isinstance(e, Cond):
        return BRANCH + expr_cost(e.c) + max(expr_cost(e.a), expr_cost(e.b))
    if isinstance(e, Op):
        w = _wi(e.width)
        if e.op in ("zext", "sext"):
            own = 0
        elif e.op == "mul":
            own = w * MUL
        elif e.op in ("divu", "divs", "modu", "mods"):
            own = w * DIV
        elif e.op == "redxor":
            own = 1 + max(0, (e.args[0].width - 1).bit_length())
        elif e.op in ("logand", "logor"):
            own = w + BRANCH
        else:
            own = w
        return own + sum(expr_cost(a) for a in e.args)
    return 1


def stmt_cost(stmts) -> int:
    total = 0
    for s in stmts:
        if isinstance(s, Assign):
            total += _wi(s.lhs.width) + expr_cost(s.rhs) + sum(expr_cost(x) for x in s.lhs.exprs())
        elif isinstance(s, If):
            total += BRANCH + expr_cost(s.cond) + max(stmt_cost(s.then), stmt_cost(s.els))
        elif isinstance(s, Case):
            total += expr_cost(s.expr)
            best = 0
            for labels, body in s.items:
                n = len(labels) if labels else 1
                best = max(best, n * (1 + BRANCH) + stmt_cost(body))
            total += best
    return total


def process_cost(p) -> int:
    return max(1, stmt_cost(p.stmts))
