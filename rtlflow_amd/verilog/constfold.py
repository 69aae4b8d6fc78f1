"""Constant folding of IR expressions (the V3Const role for constant subtrees).

Semantics are the two-state Verilog integer semantics the device op library
implements word-exactly: results are reduced modulo 2**width, shifts by at
least the operand width give 0 (or all sign bits for ``sra``), division by zero
gives 0 (include/rf_verilated.h:1856-1861 VL_DIV_III / VL_MODDIV_III).
"""
from __future__ import annotations

from ..ir import Concat, Cond, Const, Expr, MemRd, Op, Ref, Sel, mask


def _signed(v: int, w: int) -> int:
    return v - (1 << w) if w and (v >> (w - 1)) & 1 else v


def eval_op(op: str, vals, widths, width: int) -> int:
    m = mask(width)
    if op == "not":
        return ~vals[0] & m
    if op == "neg":
        return -vals[0] & m
    if op == "zext":
        return vals[0]
    if op == "sext":
        return _signed(vals[0], widths[0]) & m
    if op == "redand":
        return int(vals[0] == mask(widths[0]))
    if op == "redor":
        return int(vals[0] != 0)
    if op == "redxor":
        return bin(vals[0]).count("1") & 1
    if op == "lognot":
        return int(vals[0] == 0)
    a = vals[0]
    b = vals[1]
    if op == "and":
        return a & b
    if op == "or":
        return a | b
    if op == "xor":
        return a ^ b
    if op == "xnor":
        return ~(a ^ b) & m
    if op == "add":
        return (a + b) & m
    if op == "sub":
        return (a - b) & m
    if op == "mul":
        return (a * b) & m
    if op == "divu":
        return 0 if b == 0 else a // b
    if op == "modu":
        return 0 if b == 0 else a % b
    if op in ("divs", "mods"):
        if b == 0:
            return 0
        sa, sb = _signed(a, width), _signed(b, width)
        q = abs(sa) // abs(sb)
        if (sa < 0) != (sb < 0):
            q = -q
        if op == "divs":
            return q & m
        return (sa - q * sb) & m
    if op == "eq":
        return int(a == b)
    if op == "ne":
        return int(a != b)
    if op == "ltu":
        return int(a < b)
    if op == "leu":
        return int(a <= b)
    if op == "gtu":
        return int(a > b)
    if op == "geu":
        return int(a >= b)
    wa = widths[0]
    if op in ("lts", "les", "gts", "ges"):
        sa, sb = _signed(a, wa), _signed(b, wa)
        return int({"lts": sa < sb, "les": sa <= sb, "gts": sa > sb, "ges": sa >= sb}[op])
    if op == "logand":
        return int(a != 0 and b != 0)
    if op == "logor":
        return int(a != 0 or b != 0)
    if op == "shl":
        return 0 if b >= width else (a << b) & m
    if op == "shr":
        return 0 if b >= width else a >> b
    if op == "sra":
        sa = _signed(a, width)
        return (sa >> min(b, width)) & m
    raise ValueError(f"unknown op {op}")


def fold_expr(e: Expr) -> Expr:
    if isinstance(e, (Const, Ref)):
        return e
    if isinstance(e, MemRd):
        e.idx = fold_expr(e.idx)
        return e
    if isinstance(e, Sel):
        a = fold_expr(e.a)
        lsb = e.lsb if isinstance(e.lsb, int) else fold_expr(e.lsb)
        if isinstance(lsb, Const):
            lsb = lsb.value
        if isinstance(lsb, int):
            if lsb == 0 and e.width == a.width:
                return a
            if isinstance(a, Const):
                return Const(e.width, a.value >> lsb)
            if lsb >= a.width:
                return Const(e.width, 0)
        e.a = a
        e.lsb = lsb
        return e
    if isinstance(e, Concat):
        parts = [fold_expr(p) for p in e.parts]
        if all(isinstance(p, Const) for p in parts):
            v = 0
            for p in parts:
                v = (v << p.width) | p.value
            return Const(e.width, v)
        e.parts = parts
        return e
    if isinstance(e, Cond):
        c = fold_expr(e.c)
        a = fold_expr(e.a)
        b = fold_expr(e.b)
        if isinstance(c, Const):
            return a if c.value else b
        e.c, e.a, e.b = c, a, b
        return e
    if isinstance(e, Op):
        args = [fold_expr(x) for x in e.args]
        if all(isinstance(x, Const) for x in args):
            return Const(e.width, eval_op(e.op, [x.value for x in args],
                                          [x.width for x in args], e.width))
        e.args = args
        return e
    return e
