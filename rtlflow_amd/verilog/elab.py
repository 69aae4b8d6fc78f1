"""Elaborator: parsed Verilog modules -> width-exact design IR.

Expression sizing and signedness follow IEEE 1364-2005 section 5.4/5.5, the
rules Verilator's V3Width implements (reference: src/V3Width.cpp is not part of
this repository's scope; the rules are the language standard's):

1. An expression's size is the max of the sizes of its context-determined
   operands and, at an assignment, the size of the target.
2. Its signedness depends on its operands only (signed iff all
   context-determined operands are signed).
3. Size and type are propagated down to context-determined operands; leaves are
   zero- or sign-extended (sign-extended only if the propagated type is signed).

Self-determined operands (comparison operands as a pair, shift amounts,
reduction and logical operands, concatenation parts, indices, the condition of
?:) start a fresh context.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

from .. import ir
from ..ir import (Assign, Case, Concat, Cond, Const, Design, If, Lhs, MemRd, Op, Process,
                  Ref, Sel, Var, mask)
from .parser import (Always, ContAssign, Decl, ECall, ECond, EConcat, EBinary, EId, EIndex,
                     EIndexedRange, ENum, ERange, ERepl, EUnary, Initial, Instance, Module,
                     Param, SAssign, SBlock, SCase, SFor, SIf, SNull, VerilogError,
                     parse_verilog)

_ARITH = {"+": "add", "-": "sub", "*": "mul", "&": "and", "|": "or", "^": "xor", "~^": "xnor",
          "/": "div", "%": "mod"}
_CMP = {"==": "eq", "!=": "ne", "===": "eq", "!==": "ne", "<": "lt", "<=": "le", ">": "gt",
        ">=": "ge"}
_SHIFT = {"<<", ">>", "<<<", ">>>"}


class _Const:
    """A parameter / genvar value: (value, width, signed)."""
    __slots__ = ("value", "width", "signed")

    def __init__(self, value: int, width: int, signed: bool):
        self.width = width
        self.signed = signed
        self.value = value & mask(width)


def fold(e: ir.Expr) -> ir.Expr:
    """Constant-fold an IR expression bottom-up (V3Const's role, restricted to
    fully constant subtrees)."""
    from .constfold import fold_expr
    return fold_expr(e)


class Elaborator:
    def __init__(self, modules: Dict[str, Module], top: str, params: Optional[Dict[str, int]] = None):
        self.modules = modules
        self.top = top
        self.vars: Dict[str, Var] = {}
        self.var_list: List[Var] = []
        self.processes: List[Process] = []
        self.consts: Dict[str, _Const] = {}
        self.override = params or {}
        self.inputs: List[Var] = []
        self.outputs: List[Var] = []

    # ------------------------------------------------------------------ utils
    def err(self, node, msg):
        raise VerilogError(f"line {getattr(node, 'line', '?')}: {msg}")

    def const_value(self, e) -> _Const:
        """Evaluate a constant expression (parameters, ranges, loop bounds)."""
        w, s = self.size(e)
        x = fold(self.lower(e, w, s))
        if not isinstance(x, Const):
            self.err(e, "expression is not constant")
        return _Const(x.value, w, s)

    def const_int(self, e) -> int:
        c = self.const_value(e)
        v = c.value
        if c.signed and c.width > 0 and (v >> (c.width - 1)) & 1:
            v -= 1 << c.width
        return v

    # ------------------------------------------------------------ elaborate
    def run(self) -> Design:
        mod = self.modules.get(self.top)
        if mod is None:
            raise VerilogError(f"top module {self.top!r} not found")
        # parameters first (in order), then declarations, then logic
        for it in mod.items:
            if isinstance(it, Param):
                if not it.local and it.name in self.override:
                    val = ENum(it.line, int(self.override[it.name]), None, True)
                else:
                    val = it.value
                c = self.const_value(val)
                if it.msb is not None:
                    w = abs(self.const_int(it.msb) - self.const_int(it.lsb)) + 1
                    c = _Const(c.value, w, it.signed)
                self.consts[it.name] = c
        for it in mod.items:
            if isinstance(it, Decl):
                self.declare(it)
        for it in mod.items:
            if isinstance(it, ContAssign):
                lhs_list = self.lvalues(it.lhs)
                rhs = it.rhs
                stmts = self.assign_stmts(lhs_list, rhs, nb=False, node=it)
                self.processes.append(Process("comb", stmts, name=f"assign@{it.line}", line=it.line))
            elif isinstance(it, Always):
                if it.kind == "seq":
                    edges = []
                    for edge, cname in it.edges:
                        if cname not in self.vars:
                            self.err(it, f"unknown clock {cname}")
                        edges.append((edge, self.vars[cname]))
                    stmts = self.stmt(it.body, {})
                    self.processes.append(Process("seq", stmts, edges, name=f"always_ff@{it.line}",
                                                  line=it.line))
                else:
                    stmts = self.stmt(it.body, {})
                    self.processes.append(Process("comb", stmts, name=f"always_comb@{it.line}",
                                                  line=it.line))
            elif isinstance(it, Initial):
                stmts = self.stmt(it.body, {})
                self.processes.append(Process("initial", stmts, name=f"initial@{it.line}",
                                              line=it.line))
            elif isinstance(it, Instance):
                self.err(it, "module instantiation is not supported; flatten the design")
        for it in mod.items:
            if isinstance(it, Decl) and it.init is not None:
                v = self.vars[it.name]
                if it.kind == "wire" or (it.kind == "output" and it.net != "reg"):
                    # net declaration assignment: a continuous assign
                    stmts = self.assign_stmts([Lhs(v)], it.init, nb=False, node=it)
                    self.processes.append(Process("comb", stmts, name=f"assign@{it.line}",
                                                  line=it.line))
                    continue
                w, s = self.size(it.init)
                W = max(w, v.width)
                rhs = self.lower(it.init, W, s)
                if W > v.width:
                    rhs = Sel(rhs, 0, v.width)
                rhs = fold(rhs)
                if isinstance(rhs, Const) and not v.is_mem:
                    v.init = rhs.value
                self.processes.insert(0, Process("initial", [Assign(Lhs(v), rhs, False)],
                                                 name=f"init {v.name}", line=it.line))
        d = Design(self.top, self.var_list, self.processes, self.inputs, self.outputs)
        d.finalize()
        return d

    def declare(self, d: Decl):
        if d.name in self.vars:
            v = self.vars[d.name]
            # non-ANSI: "output x; reg [3:0] x;" refines the declaration
            if d.msb is not None:
                v.width = abs(self.const_int(d.msb) - self.const_int(d.lsb)) + 1
            v.signed = v.signed or d.signed
            if d.kind in ("input", "output"):
                v.kind = d.kind
                v.is_port = True
                (self.inputs if d.kind == "input" else self.outputs).append(v)
            return
        width = 1
        if d.msb is not None:
            msb = self.const_int(d.msb)
            lsb = self.const_int(d.lsb)
            if lsb != 0 and msb != 0 and min(msb, lsb) != 0:
                self.err(d, "packed ranges must end at bit 0")
            width = abs(msb - lsb) + 1
        depth = None
        lo = 0
        if d.array is not None:
            a = self.const_int(d.array[0])
            b = self.const_int(d.array[1])
            depth = abs(a - b) + 1
            lo = min(a, b)
        kind = d.kind
        if kind == "inout":
            self.err(d, "inout ports are not supported")
        v = Var(d.name, width, d.signed, kind if kind in ("input", "output") else
                ("reg" if kind in ("reg", "integer") else "wire"), depth, lo)
        self.vars[d.name] = v
        self.var_list.append(v)
        if kind == "input":
            self.inputs.append(v)
        elif kind == "output":
            self.outputs.append(v)

    # --------------------------------------------------------------- sizing
    def size(self, e) -> Tuple[int, bool]:
        """Self-determined (width, signed) of a parse-tree expression."""
        if isinstance(e, ENum):
            return (e.width if e.width is not None else 32), e.signed
        if isinstance(e, EId):
            if e.name in self.consts:
                c = self.consts[e.name]
                return c.width, c.signed
            v = self.vars.get(e.name)
            if v is None:
                self.err(e, f"unknown identifier {e.name}")
            if v.is_mem:
                self.err(e, f"memory {e.name} used without index")
            return v.width, v.signed
        if isinstance(e, EIndex):
            if isinstance(e.base, EId) and e.base.name in self.vars and self.vars[e.base.name].is_mem:
                v = self.vars[e.base.name]
                return v.width, v.signed
            return 1, False
        if isinstance(e, ERange):
            return abs(self.const_int(e.msb) - self.const_int(e.lsb)) + 1, False
        if isinstance(e, EIndexedRange):
            return self.const_int(e.width), False
        if isinstance(e, EConcat):
            return sum(self.size(p)[0] for p in e.parts), False
        if isinstance(e, ERepl):
            return self.const_int(e.count) * sum(self.size(p)[0] for p in e.parts), False
        if isinstance(e, EUnary):
            if e.op in ("~", "-", "+"):
                return self.size(e.a)
            return 1, False
        if isinstance(e, EBinary):
            if e.op in _CMP or e.op in ("&&", "||"):
                return 1, False
            if e.op in _SHIFT or e.op == "**":
                return self.size(e.a)
            wa, sa = self.size(e.a)
            wb, sb = self.size(e.b)
            return max(wa, wb), sa and sb
        if isinstance(e, ECond):
            wa, sa = self.size(e.a)
            wb, sb = self.size(e.b)
            return max(wa, wb), sa and sb
        if isinstance(e, ECall):
            if e.name == "$signed":
                return self.size(e.args[0])[0], True
            if e.name == "$unsigned":
                return self.size(e.args[0])[0], False
            if e.name == "$clog2":
                return 32, True
            self.err(e, f"unsupported function {e.name}")
        self.err(e, f"unsupported expression {type(e).__name__}")

    # -------------------------------------------------------------- lowering
    def extend(self, x: ir.Expr, W: int, signed: bool) -> ir.Expr:
        if x.width == W:
            return x
        if x.width > W:
            return Sel(x, 0, W)
        return Op("sext" if signed else "zext", [x], W)

    def lower_self(self, e) -> ir.Expr:
        w, s = self.size(e)
        return self.lower(e, w, s)

    def lower(self, e, W: int, S: bool) -> ir.Expr:
        """Lower ``e`` in a context of width W and signedness S."""
        if isinstance(e, ENum):
            w = e.width if e.width is not None else 32
            return self.extend(Const(w, e.value), W, S)
        if isinstance(e, EId):
            if e.name in self.consts:
                c = self.consts[e.name]
                return self.extend(Const(c.width, c.value), W, S)
            v = self.vars[e.name]
            return self.extend(Ref(v), W, S)
        if isinstance(e, (EIndex, ERange, EIndexedRange)):
            w, _ = self.size(e)
            return self.extend(self.select(e), W, S)
        if isinstance(e, EConcat):
            parts = [self.lower_self(p) for p in e.parts]
            x = parts[0] if len(parts) == 1 else Concat(parts)
            return self.extend(x, W, S)
        if isinstance(e, ERepl):
            n = self.const_int(e.count)
            parts = [self.lower_self(p) for p in e.parts]
            if n <= 0:
                self.err(e, "replication count must be positive")
            x = Concat(parts * n) if (len(parts) * n) > 1 else parts[0]
            return self.extend(x, W, S)
        if isinstance(e, EUnary):
            if e.op == "+":
                return self.lower(e.a, W, S)
            if e.op == "~":
                return Op("not", [self.lower(e.a, W, S)], W)
            if e.op == "-":
                return Op("neg", [self.lower(e.a, W, S)], W)
            a = self.lower_self(e.a)
            if e.op == "!":
                r = Op("lognot", [a], 1)
            elif e.op == "&":
                r = Op("redand", [a], 1)
            elif e.op == "|":
                r = Op("redor", [a], 1)
            elif e.op == "^":
                r = Op("redxor", [a], 1)
            elif e.op == "~&":
                r = Op("not", [Op("redand", [a], 1)], 1)
            elif e.op == "~|":
                r = Op("not", [Op("redor", [a], 1)], 1)
            elif e.op == "~^":
                r = Op("not", [Op("redxor", [a], 1)], 1)
            else:
                self.err(e, f"unsupported unary {e.op}")
            return self.extend(r, W, False)
        if isinstance(e, EBinary):
            op = e.op
            if op in _ARITH:
                a = self.lower(e.a, W, S)
                b = self.lower(e.b, W, S)
                name = _ARITH[op]
                if name in ("div", "mod"):
                    name += "s" if S else "u"
                return Op(name, [a, b], W)
            if op in _CMP:
                wa, sa = self.size(e.a)
                wb, sb = self.size(e.b)
                w2 = max(wa, wb)
                s2 = sa and sb
                a = self.lower(e.a, w2, s2)
                b = self.lower(e.b, w2, s2)
                name = _CMP[op]
                if name not in ("eq", "ne"):
                    name += "s" if s2 else "u"
                return self.extend(Op(name, [a, b], 1), W, False)
            if op in ("&&", "||"):
                a = self.lower_self(e.a)
                b = self.lower_self(e.b)
                return self.extend(Op("logand" if op == "&&" else "logor", [a, b], 1), W, False)
            if op in _SHIFT:
                a = self.lower(e.a, W, S)
                b = self.lower_self(e.b)
                if op in ("<<", "<<<"):
                    name = "shl"
                elif op == ">>>" and S:
                    name = "sra"
                else:
                    name = "shr"
                return Op(name, [a, b], W)
            if op == "**":
                base = self.const_value(e.a)
                ex = self.const_int(e.b)
                return self.extend(Const(base.width, base.value ** max(ex, 0)), W, S)
            self.err(e, f"unsupported binary operator {op}")
        if isinstance(e, ECond):
            c = self.truth(self.lower_self(e.c))
            return Cond(c, self.lower(e.a, W, S), self.lower(e.b, W, S))
        if isinstance(e, ECall):
            if e.name in ("$signed", "$unsigned"):
                x = self.lower_self(e.args[0])
                return self.extend(x, W, S)
            if e.name == "$clog2":
                v = self.const_int(e.args[0])
                r = 0
                while (1 << r) < v:
                    r += 1
                return self.extend(Const(32, r), W, S)
        self.err(e, f"unsupported expression {type(e).__name__}")

    def truth(self, x: ir.Expr) -> ir.Expr:
        return x if x.width == 1 else Op("redor", [x], 1)

    def base_var(self, e) -> Var:
        if not isinstance(e, EId) or e.name not in self.vars:
            self.err(e, "select base must be a declared variable")
        return self.vars[e.name]

    def index_expr(self, e, offset: int = 0) -> ir.Expr:
        x = self.lower_self(e)
        if offset:
            x = Op("sub", [self.extend(x, max(x.width, 32), False), Const(max(x.width, 32), offset)],
                   max(x.width, 32))
        return fold(x)

    def select(self, e) -> ir.Expr:
        """Read-side bit/part select or memory element (possibly with a further select)."""
        if isinstance(e, EIndex):
            base = e.base
            if isinstance(base, EId) and base.name in self.vars and self.vars[base.name].is_mem:
                v = self.vars[base.name]
                return MemRd(v, self.index_expr(e.index, v.lo))
            src = self.select_base(base)
            idx = fold(self.lower_self(e.index))
            if isinstance(idx, Const):
                return Sel(src, idx.value, 1)
            return Sel(src, idx, 1)
        if isinstance(e, ERange):
            src = self.select_base(e.base)
            msb = self.const_int(e.msb)
            lsb = self.const_int(e.lsb)
            lo, hi = min(msb, lsb), max(msb, lsb)
            return Sel(src, lo, hi - lo + 1)
        if isinstance(e, EIndexedRange):
            src = self.select_base(e.base)
            w = self.const_int(e.width)
            start = fold(self.lower_self(e.start))
            if not e.up:
                if isinstance(start, Const):
                    return Sel(src, start.value - w + 1, w)
                sw = max(start.width, 32)
                start = Op("sub", [self.extend(start, sw, False), Const(sw, w - 1)], sw)
            if isinstance(start, Const):
                return Sel(src, start.value, w)
            return Sel(src, start, w)
        raise AssertionError

    def select_base(self, base) -> ir.Expr:
        if isinstance(base, EId):
            if base.name in self.consts:
                c = self.consts[base.name]
                return Const(c.width, c.value)
            v = self.vars.get(base.name)
            if v is None:
                self.err(base, f"unknown identifier {base.name}")
            if v.is_mem:
                self.err(base, "memory needs an element index")
            return Ref(v)
        if isinstance(base, EIndex):
            # memory element then bit select
            return self.select(base)
        if isinstance(base, (EConcat, ERepl)):
            return self.lower_self(base)
        self.err(base, "unsupported select base")

    # ------------------------------------------------------------ statements
    def lvalues(self, e) -> List[Lhs]:
        """Targets of an assignment, most significant first (for concatenations)."""
        if isinstance(e, EConcat):
            out = []
            for p in e.parts:
                out.extend(self.lvalues(p))
            return out
        if isinstance(e, EId):
            v = self.base_var(e)
            if v.is_mem:
                self.err(e, "cannot assign a whole memory")
            return [Lhs(v)]
        if isinstance(e, EIndex):
            base = e.base
            if isinstance(base, EId) and self.base_var(base).is_mem:
                v = self.base_var(base)
                return [Lhs(v, idx=self.index_expr(e.index, v.lo))]
            if isinstance(base, EIndex) and isinstance(base.base, EId) and self.base_var(base.base).is_mem:
                v = self.base_var(base.base)
                idx = self.index_expr(base.index, v.lo)
                b = fold(self.lower_self(e.index))
                return [Lhs(v, idx=idx, lsb=(b.value if isinstance(b, Const) else b), width=1)]
            v = self.base_var(base)
            b = fold(self.lower_self(e.index))
            return [Lhs(v, lsb=(b.value if isinstance(b, Const) else b), width=1)]
        if isinstance(e, (ERange, EIndexedRange)):
            base = e.base
            idx = None
            if isinstance(base, EIndex):
                v = self.base_var(base.base)
                if not v.is_mem:
                    self.err(e, "unsupported nested select target")
                idx = self.index_expr(base.index, v.lo)
            else:
                v = self.base_var(base)
            if isinstance(e, ERange):
                msb = self.const_int(e.msb)
                lsb = self.const_int(e.lsb)
                lo, hi = min(msb, lsb), max(msb, lsb)
                return [Lhs(v, idx=idx, lsb=lo, width=hi - lo + 1)]
            w = self.const_int(e.width)
            start = fold(self.lower_self(e.start))
            if not e.up:
                if isinstance(start, Const):
                    return [Lhs(v, idx=idx, lsb=start.value - w + 1, width=w)]
                sw = max(start.width, 32)
                start = Op("sub", [self.extend(start, sw, False), Const(sw, w - 1)], sw)
            return [Lhs(v, idx=idx, lsb=(start.value if isinstance(start, Const) else start), width=w)]
        self.err(e, "unsupported assignment target")

    def assign_stmts(self, lhs_list: List[Lhs], rhs, nb: bool, node) -> List[ir.Stmt]:
        L = sum(l.width for l in lhs_list)
        w, s = self.size(rhs)
        W = max(L, w)
        r = fold(self.lower(rhs, W, s))
        if W > L:
            r = fold(Sel(r, 0, L))
        if len(lhs_list) == 1:
            return [Assign(lhs_list[0], r, nb)]
        out = []
        pos = L
        for l in lhs_list:
            pos -= l.width
            out.append(Assign(l, fold(Sel(r, pos, l.width)), nb))
        return out

    def stmt(self, s, loopvars: Dict[str, int]) -> List[ir.Stmt]:
        if isinstance(s, SBlock):
            out = []
            for x in s.stmts:
                out.extend(self.stmt(x, loopvars))
            return out
        if isinstance(s, SNull):
            return []
        if isinstance(s, SAssign):
            return self.assign_stmts(self.lvalues(s.lhs), s.rhs, s.nonblocking, s)
        if isinstance(s, SIf):
            c = fold(self.truth(self.lower_self(s.cond)))
            th = self.stmt(s.then, loopvars) if s.then is not None else []
            el = self.stmt(s.els, loopvars) if s.els is not None else []
            if isinstance(c, Const):
                return th if c.value else el
            return [If(c, th, el)]
        if isinstance(s, SCase):
            return self.case(s, loopvars)
        if isinstance(s, SFor):
            return self.unroll(s, loopvars)
        self.err(s, f"unsupported statement {type(s).__name__}")

    def case(self, s: SCase, loopvars) -> List[ir.Stmt]:
        sizes = [self.size(s.expr)]
        for labels, _ in s.items:
            if labels:
                sizes.extend(self.size(l) for l in labels)
        W = max(w for w, _ in sizes)
        S = all(sg for _, sg in sizes)
        sel = fold(self.lower(s.expr, W, S))
        items = []
        for labels, body in s.items:
            stmts = self.stmt(body, loopvars) if body is not None else []
            if labels is None:
                items.append((None, stmts))
                continue
            lab = []
            for l in labels:
                care = mask(W)
                if s.kind in ("casez", "casex") and isinstance(l, ENum):
                    care = self.care_mask(l, W, s.kind)
                x = fold(self.lower(l, W, S))
                if not isinstance(x, Const):
                    self.err(l, "case labels must be constant")
                lab.append((x.value & care, care))
            items.append((lab, stmts))
        return [Case(sel, items)]

    def care_mask(self, num: ENum, W: int, kind: str) -> int:
        """Bits of a casez/casex label that are compared (z/? and, for casex, x
        are don't-care). Recomputed from the literal text kept by the lexer."""
        dc = num.zmask | (num.xmask if kind == "casex" else 0)
        return mask(W) & ~dc

    def unroll(self, s: SFor, loopvars: Dict[str, int]) -> List[ir.Stmt]:
        out = []
        saved = self.consts.get(s.var)
        self.consts[s.var] = _Const(self.const_int(s.init), 32, True)
        n = 0
        try:
            while self.const_int(s.cond):
                out.extend(self.stmt(s.body, loopvars))
                self.consts[s.var] = _Const(self.const_int(s.step), 32, True)
                n += 1
                if n > 1_000_000:
                    self.err(s, "for loop does not terminate")
        finally:
            if saved is None:
                del self.consts[s.var]
            else:
                self.consts[s.var] = saved
        return out


def elaborate(src_or_modules, top: Optional[str] = None, params: Optional[Dict[str, int]] = None) -> Design:
    if isinstance(src_or_modules, str):
        mods = parse_verilog(src_or_modules)
    else:
        mods = src_or_modules
    table = {m.name: m for m in mods}
    if top is None:
        top = mods[-1].name
    return Elaborator(table, top, params).run()
