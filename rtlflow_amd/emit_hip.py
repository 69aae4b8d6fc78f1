"""HIP code generator for gfx950 (the rewrite of the reference's V3EmitC
RTLflow emitter, src/V3EmitC.cpp).

For a compiled design it writes one HIP source holding:

* one ``__global__`` kernel per mtask, one lane (stimulus) per thread
  (reference: ``__Vmtask__<id>``, src/V3EmitC.cpp:1637-1662) -- lanes masked by
  ``done[]`` and, when the design has change-detected logic, by ``change[]``
  return early exactly as the reference's kernels do (:1654);
* ``rf_k_init``: initial blocks, clock latch and settle for each lane
  (reference: host OpenMP ``_eval_initial`` + ``_eval_settle`` kernel,
  src/V3Clock.cpp:228-270, src/V3EmitC.cpp:1728-1743, :5466-5470);
* ``rf_k_eval_end``: ``__Vclklast__`` latch (``_last_assign``,
  src/V3Clock.cpp:214-226) fused with the change request
  (``_change_request``, src/V3Changed.cpp:263-275) and a wavefront-reduced OR of
  the per-lane change flags into one device word (the reference's Taskflow
  ``cudaFlow::reduce`` node, src/V3EmitC.cpp:5395-5398);
* ``rf_k_stim``: the on-device stimulus for one cycle (replaces the host writes
  through ``RTLflow::get`` pointers into managed memory, src/V3EmitC.cpp:5339-5350);
* ``rf_k_fused``: a per-lane kernel running whole clock cycles (inputs, clk=0
  eval, clk=1 eval with the settle loop on device, output hash) with design
  state in registers;
* ``rf_k_persistent``: the batched form -- a persistent kernel whose lanes
  take stimuli from a device work queue and refill on finish;
* the static ``rf_design`` descriptor table (include/rf_design.h).

All state lives in the lane-minor SoA layout of layout.py.
"""
from __future__ import annotations

import hashlib
from typing import Dict, List, Optional, Set, Tuple

from .ir import (Assign, Case, Concat, Cond, Const, Design, Expr, If, Lhs, MemRd, Op, Process,
                 Ref, Sel, Stmt, Var, mask, words_of)
from .layout import Layout, class_of
from .order import CompileError

CTY = {"C": "uint8_t", "S": "uint16_t", "I": "uint32_t", "Q": "uint64_t"}
ARR = {"C": "sC", "S": "sS", "I": "sI", "Q": "sQ"}
EBYTES = {"C": 1, "S": 2, "I": 4, "Q": 8}


def cty(w: int) -> str:
    return "uint32_t" if w <= 32 else "uint64_t"


def lit(v: int, w: int) -> str:
    if w <= 32:
        return f"0x{v:x}u"
    return f"0x{v:x}ull"


def is_wide(w: int) -> bool:
    return w > 64


class Gen:
    """Statement/expression emitter for one kernel body.

    ``local``: vars kept in registers (C locals named v<id>), all others are
    accessed in SoA memory on every use. ``avail``: register vars whose local
    currently holds the variable's value. ``dirty``: register vars modified
    since their last store."""

    def __init__(self, comp, in_kernel: bool = True):
        self.comp = comp
        self.lines: List[str] = []
        self.ind = 1
        self.ntmp = 0
        self.local: Set[Var] = set()
        self.avail: Set[Var] = set()
        self.dirty: Set[Var] = set()
        self.decl: Set[Var] = set()
        self.nb: Dict[Var, str] = {}      # var -> NBA temp name inside a seq process
        self.mem_nb: List[Tuple[str, MemRd]] = []
        self.in_kernel = in_kernel
        # 1-bit register vars packed into bit slots of 32-bit registers:
        # var -> (register name, bit)
        self.packed: Dict[Var, Tuple[str, int]] = {}
        self.pack_regs: Set[str] = set()
        # vars whose NBA writes must go through a temporary (read by the
        # same clocked group); None = every NBA target
        self.nb_needed: Optional[Set[Var]] = None
        self.write_through = False
        # small memories held in LDS by whole-cycle kernels: var -> array name
        # (element word k of this thread at name[k][threadIdx.x])
        self.lds: Dict[Var, str] = {}
        # register variables kept in LDS instead (the state of designs too
        # large for the register file): narrow var -> lvalue of this
        # thread's cell; wide var -> initialiser of its word pointer
        self.lds_vars: Dict[Var, str] = {}

    # ------------------------------------------------------------ output
    def emit(self, s: str):
        self.lines.append("    " * self.ind + s)

    def tmp(self, prefix="t") -> str:
        self.ntmp += 1
        return f"{prefix}{self.ntmp}"

    # ------------------------------------------------------------ storage
    def name(self, v: Var) -> str:
        return f"v{v.id}"

    def get(self, v: Var) -> str:
        """Rvalue of a narrow register var."""
        p = self.packed.get(v)
        if p is not None:
            return f"rf_bget({p[0]}, {p[1]})"
        cell = self.lds_vars.get(v)
        if cell is not None:
            return f"(({cty(v.width)}){cell})"
        return self.name(v)

    def set(self, v: Var, expr: str) -> str:
        """Statement assigning a narrow register var."""
        p = self.packed.get(v)
        if p is not None:
            return f"{p[0]} = rf_bset({p[0]}, {p[1]}, {expr});"
        cell = self.lds_vars.get(v)
        if cell is not None:
            return f"{cell} = ({CTY[class_of(v)]})({expr});"
        return f"{self.name(v)} = {expr};"

    def soa(self, v: Var, word: int = 0) -> str:
        loc = self.comp.layout.loc[v]
        c = class_of(v)
        return f"{ARR[c]}[(size_t){loc + word} * NP]"

    def declare(self, v: Var):
        if v in self.decl:
            return
        self.decl.add(v)
        p = self.packed.get(v)
        if p is not None:
            if p[0] not in self.pack_regs:
                self.pack_regs.add(p[0])
                self.emit_top(f"uint32_t {p[0]} = 0u;")
            return
        cell = self.lds_vars.get(v)
        if cell is not None:
            if is_wide(v.width):
                self.emit_top(f"uint32_t* const {self.name(v)} = {cell};")
            else:
                self.emit_top(f"{cell} = 0;")
            return
        if is_wide(v.width):
            self.emit_top(f"uint32_t {self.name(v)}[{v.words}];")
        else:
            self.emit_top(f"{cty(v.width)} {self.name(v)} = 0;")

    def emit_top(self, s: str):
        self.top_lines.append("    " + s)

    def load(self, v: Var):
        """Make the local of register var ``v`` hold its stored value."""
        n = self.name(v)
        if is_wide(v.width):
            for w in range(v.words):
                self.emit(f"{n}[{w}] = {self.soa(v, w)};")
        else:
            self.emit(self.set(v, f"({cty(v.width)}){self.soa(v)}"))

    def store(self, v: Var):
        n = self.name(v)
        c = class_of(v)
        if is_wide(v.width):
            for w in range(v.words):
                self.emit(f"{self.soa(v, w)} = {n}[{w}];")
        else:
            self.emit(f"{self.soa(v)} = ({CTY[c]}){self.get(v)};")

    def ensure(self, v: Var):
        """Before a read of ``v`` (non-memory)."""
        if v in self.local:
            self.declare(v)
            if v not in self.avail:
                self.load(v)
                self.avail.add(v)

    # --------------------------------------------------------- var reads
    def read_narrow(self, v: Var) -> str:
        if v in self.local:
            return self.get(v)
        return f"(({cty(v.width)}){self.soa(v)})"

    def read_wide(self, v: Var) -> str:
        if v in self.local:
            return self.name(v)
        t = self.tmp()
        self.emit(f"uint32_t {t}[{v.words}];")
        for w in range(v.words):
            self.emit(f"{t}[{w}] = {self.soa(v, w)};")
        return t

    # --------------------------------------------------------- memories
    def mem_elem(self, m: Var, idx: str, word: int = 0) -> str:
        if m in self.lds:
            return f"{self.lds[m]}[(size_t)({idx}) * {m.words} + {word}][threadIdx.x]"
        lay = self.comp.layout
        if m in lay.lane_major:
            c = class_of(m)
            return f"mb{m.id}[(size_t)({idx}) * {m.words} + {word}]"
        loc = lay.loc[m]
        return f"{ARR[class_of(m)]}[((size_t){loc} + (size_t)({idx}) * {m.words} + {word}) * NP]"

    def idx_tmp(self, e: Expr) -> str:
        if is_wide(e.width):
            a = self.exw(e)
            t = self.tmp("ix")
            self.emit(f"const uint64_t {t} = rfw_amount({a}, {e.width});")
            return t
        t = self.tmp("ix")
        self.emit(f"const uint64_t {t} = (uint64_t)({self.ex(e)});")
        return t

    # ------------------------------------------------- narrow expressions
    def amount(self, e: Expr) -> str:
        if is_wide(e.width):
            return f"rfw_amount({self.exw(e)}, {e.width})"
        return f"(uint64_t)({self.ex(e)})"

    def ex(self, e: Expr) -> str:
        """C expression (clean, type cty(width)) for a narrow expression."""
        w = e.width
        assert w <= 64, e
        if isinstance(e, Const):
            return lit(e.value, w)
        if isinstance(e, Ref):
            self.ensure(e.var)
            return self.read_narrow(e.var)
        if isinstance(e, MemRd):
            m = e.var
            ix = self.idx_tmp(e.idx)
            return f"({ix} < {m.depth}u ? ({cty(w)}){self.mem_elem(m, ix)} : ({cty(w)})0)"
        if isinstance(e, Sel):
            a = e.a
            if is_wide(a.width):
                arr = self.exw(a)
                lsb = str(e.lsb) if isinstance(e.lsb, int) else self.amount(e.lsb)
                return f"(({cty(w)})rfw_get64({arr}, {a.width}, {lsb}, {w}))"
            src = self.ex(a)
            if isinstance(e.lsb, int):
                if e.lsb == 0:
                    body = src
                else:
                    body = f"({src} >> {e.lsb})"
                m = mask(w)
                if w < a.width:
                    body = f"({body} & {lit(m, max(a.width, w))})"
                return f"(({cty(w)}){body})"
            t = self.tmp("sh")
            self.emit(f"const uint64_t {t} = {self.amount(e.lsb)};")
            return (f"(({cty(w)})({t} >= {a.width}u ? 0ull : ((uint64_t)({src}) >> {t}) & "
                    f"{lit(mask(w), 64)}))")
        if isinstance(e, Concat):
            parts = []
            pos = w
            for p in e.parts:
                pos -= p.width
                x = self.ex(p)
                if pos:
                    parts.append(f"(({cty(w)})({x}) << {pos})")
                else:
                    parts.append(f"(({cty(w)})({x}))")
            return "(" + " | ".join(parts) + ")"
        if isinstance(e, Cond):
            c = self.ex(e.c)
            a = self.ex(e.a)
            b = self.ex(e.b)
            return f"({c} ? {a} : {b})"
        if isinstance(e, Op):
            return self.ex_op(e)
        raise CompileError(f"cannot emit {e!r}")

    def msk(self, s: str, w: int) -> str:
        if w in (32, 64):
            return s
        return f"(({s}) & {lit(mask(w), w)})"

    def ex_op(self, e: Op) -> str:
        op, w = e.op, e.width
        T = cty(w)
        args = e.args
        if op in ("eq", "ne", "ltu", "leu", "gtu", "geu", "lts", "les", "gts", "ges"):
            wa = args[0].width
            if is_wide(wa):
                a = self.exw(args[0])
                b = self.exw(args[1])
                if op == "eq":
                    return f"rfw_eq({a}, {b}, {wa})"
                if op == "ne":
                    return f"(rfw_eq({a}, {b}, {wa}) ^ 1u)"
                fn = "rfw_cmps" if op in ("lts", "les", "gts", "ges") else "rfw_cmpu"
                rel = {"lt": "<", "le": "<=", "gt": ">", "ge": ">="}[op[:2]]
                return f"((uint32_t)({fn}({a}, {b}, {wa}) {rel} 0))"
            a = self.ex(args[0])
            b = self.ex(args[1])
            rel = {"eq": "==", "ne": "!=", "lt": "<", "le": "<=", "gt": ">", "ge": ">="}[
                op if op in ("eq", "ne") else op[:2]]
            if op in ("lts", "les", "gts", "ges"):
                sx = "rf_sx32" if wa <= 32 else "rf_sx64"
                return f"((uint32_t)({sx}({a}, {wa}) {rel} {sx}({b}, {wa})))"
            return f"((uint32_t)({a} {rel} {b}))"
        if op in ("redand", "redor", "redxor", "lognot"):
            wa = args[0].width
            if is_wide(wa):
                a = self.exw(args[0])
                if op == "lognot":
                    return f"(rfw_redor({a}, {wa}) ^ 1u)"
                return f"rfw_{op}({a}, {wa})"
            a = self.ex(args[0])
            if op == "redand":
                return f"((uint32_t)({a} == {lit(mask(wa), wa)}))"
            if op == "redor":
                return f"((uint32_t)({a} != 0))"
            if op == "lognot":
                return f"((uint32_t)({a} == 0))"
            return f"rf_parity{32 if wa <= 32 else 64}({a})"
        if op in ("logand", "logor"):
            xs = []
            for x in args:
                if is_wide(x.width):
                    xs.append(f"rfw_redor({self.exw(x)}, {x.width})")
                else:
                    xs.append(f"({self.ex(x)} != 0)")
            rel = "&&" if op == "logand" else "||"
            return f"((uint32_t)({xs[0]} {rel} {xs[1]}))"
        if op == "zext":
            return f"(({T})({self.ex(args[0])}))"
        if op == "sext":
            wa = args[0].width
            a = self.ex(args[0])
            if w <= 32:
                return self.msk(f"(uint32_t)rf_sx32({a}, {wa})", w)
            return self.msk(f"(uint64_t)rf_sx64({a}, {wa})", w)
        if op in ("shl", "shr", "sra"):
            a = self.ex(args[0])
            amt = self.amount(args[1])
            bits = 32 if w <= 32 else 64
            return f"rf_{op}{bits}({a}, {amt}, {w})"
        if op == "not":
            return self.msk(f"~{self.ex(args[0])}", w)
        if op == "neg":
            return self.msk(f"(({T})0 - {self.ex(args[0])})", w)
        a = self.ex(args[0])
        b = self.ex(args[1])
        if op == "and":
            return f"({a} & {b})"
        if op == "or":
            return f"({a} | {b})"
        if op == "xor":
            return f"({a} ^ {b})"
        if op == "xnor":
            return self.msk(f"~({a} ^ {b})", w)
        if op == "add":
            return self.msk(f"({a} + {b})", w)
        if op == "sub":
            return self.msk(f"({a} - {b})", w)
        if op == "mul":
            return self.msk(f"({a} * {b})", w)
        bits = 32 if w <= 32 else 64
        if op in ("divu", "modu"):
            return f"rf_{op}{bits}({a}, {b})"
        if op in ("divs", "mods"):
            return f"rf_{op}{bits}({a}, {b}, {w})"
        raise CompileError(f"operator {op} not supported")

    # --------------------------------------------------- wide expressions
    def exw(self, e: Expr, dirty_ok: bool = False) -> str:
        """Name of a uint32_t array holding the (clean) wide value of ``e``.
        A word-aligned slice of another wide value is a view into that
        value's words instead of a copy, when it is clean (its top bit is a
        word boundary or the source's top) or ``dirty_ok`` (the consumer
        masks the top word itself)."""
        w = e.width
        n = words_of(w)
        if isinstance(e, Ref):
            self.ensure(e.var)
            return self.read_wide(e.var)
        if isinstance(e, Sel) and isinstance(e.lsb, int) and e.lsb % 32 == 0 and is_wide(e.a.width) \
                and e.lsb + w <= e.a.width and (dirty_ok or w % 32 == 0 or e.lsb + w == e.a.width):
            a = self.exw(e.a, dirty_ok=True)
            return f"({a} + {e.lsb // 32})" if e.lsb else a
        t = self.tmp("w")
        if isinstance(e, Const):
            ws = ", ".join(f"0x{(e.value >> (32 * i)) & 0xffffffff:x}u" for i in range(n))
            self.emit(f"const uint32_t {t}[{n}] = {{{ws}}};")
            return t
        self.emit(f"uint32_t {t}[{n}];")
        if isinstance(e, MemRd):
            m = e.var
            ix = self.idx_tmp(e.idx)
            self.emit(f"if ({ix} < {m.depth}u) {{")
            for k in range(n):
                self.emit(f"    {t}[{k}] = {self.mem_elem(m, ix, k)};")
            self.emit(f"}} else {{ rfw_zero({t}, {w}); }}")
            return t
        if isinstance(e, Sel):
            a = self.exw(e.a)
            lsb = str(e.lsb) if isinstance(e.lsb, int) else self.amount(e.lsb)
            self.emit(f"rfw_extract({t}, {w}, {a}, {e.a.width}, {lsb});")
            return t
        if isinstance(e, Concat):
            self.emit(f"rfw_zero({t}, {w});")
            pos = w
            for p in e.parts:
                pos -= p.width
                if is_wide(p.width):
                    self.emit(f"rfw_putw({t}, {w}, {pos}, {self.exw(p, dirty_ok=True)}, {p.width});")
                else:
                    self.emit(f"rfw_put64({t}, {w}, {pos}, {p.width}, (uint64_t)({self.ex(p)}));")
            return t
        if isinstance(e, Cond):
            c = self.ex(e.c)
            a = self.exw(e.a)
            b = self.exw(e.b)
            self.emit(f"rfw_cond({t}, {c}, {a}, {b}, {w});")
            return t
        if isinstance(e, Op):
            op = e.op
            args = e.args
            if op in ("zext", "sext"):
                x = args[0]
                if is_wide(x.width):
                    fn = "rfw_zext" if op == "zext" else "rfw_sextw"
                    self.emit(f"{fn}({t}, {w}, {self.exw(x)}, {x.width});")
                elif op == "zext":
                    self.emit(f"rfw_from64({t}, {w}, (uint64_t)({self.ex(x)}));")
                else:
                    self.emit(f"rfw_sext64({t}, {w}, (uint64_t)({self.ex(x)}), {x.width});")
                return t
            if op in ("shl", "shr", "sra"):
                a = self.exw(args[0])
                amt = self.amount(args[1])
                self.emit(f"rfw_{op}({t}, {a}, {amt}, {w});")
                return t
            if op in ("not", "neg"):
                self.emit(f"rfw_{op}({t}, {self.exw(args[0])}, {w});")
                return t
            if op in ("and", "or", "xor", "xnor", "add", "sub", "mul"):
                a = self.exw(args[0])
                b = self.exw(args[1])
                self.emit(f"rfw_{op}({t}, {a}, {b}, {w});")
                return t
            if op in ("divu", "modu", "divs", "mods"):
                if words_of(w) > 16:
                    raise CompileError(f"division wider than 512 bits is not supported ({w})")
                a = self.exw(args[0])
                b = self.exw(args[1])
                self.emit(f"rfw_{op}({t}, {a}, {b}, {w});")
                return t
        raise CompileError(f"wide expression {e!r} not supported")

    # ------------------------------------------------------- statements
    def target(self, v: Var) -> str:
        return self.nb.get(v, self.name(v))

    def assign(self, s: Assign):
        lhs = s.lhs
        v = lhs.var
        if v.is_mem:
            self.mem_assign(s)
            return
        if v not in self.local:
            raise CompileError(f"internal: write to non-register var {v.name}")
        self.declare(v)
        dst = self.target(v)
        in_nb = v in self.nb
        if lhs.whole:
            shift_in = self._shift_in(v, s.rhs) if is_wide(v.width) and not in_nb else None
            if shift_in is not None:
                k, low = shift_in
                self.ensure(v)
                self.emit(f"rfw_shl_in({dst}, {v.width}, {k}, (uint64_t)({low}));")
            elif is_wide(v.width):
                src = self.exw(s.rhs)
                self.emit(f"rfw_copy({dst}, {src}, {v.width});")
            elif in_nb:
                self.emit(f"{dst} = {self.ex(s.rhs)};")
            else:
                self.emit(self.set(v, self.ex(s.rhs)))
        else:
            if v not in self.avail and not in_nb:
                self.ensure(v)
            lsb = str(lhs.lsb) if isinstance(lhs.lsb, int) else self.amount(lhs.lsb)
            if is_wide(v.width):
                if is_wide(lhs.width):
                    self.emit(f"rfw_putw({dst}, {v.width}, {lsb}, {self.exw(s.rhs)}, {lhs.width});")
                else:
                    self.emit(f"rfw_put64({dst}, {v.width}, {lsb}, {lhs.width}, "
                              f"(uint64_t)({self.ex(s.rhs)}));")
            else:
                cur = dst if in_nb else self.get(v)
                val = (f"({cty(v.width)})rf_ins64((uint64_t){cur}, {v.width}, {lsb}, "
                       f"{lhs.width}, (uint64_t)({self.ex(s.rhs)}))")
                self.emit(f"{dst} = {val};" if in_nb else self.set(v, val))
        if not in_nb:
            self.dirty.add(v)
            # the register now holds the variable's current value on this path
            # (a later partial write must not reload the stale stored copy)
            self.avail.add(v)
            self._write_through(v)

    def _shift_in(self, v: Var, rhs: Expr):
        """``v = {v[W-1-k:0], low}`` with ``low`` at most 64 bits: the
        shift register idiom, done in place -> (k, low's expression)."""
        if not isinstance(rhs, Concat) or len(rhs.parts) < 2:
            return None
        head = rhs.parts[0]
        rest = rhs.parts[1:]
        k = sum(p.width for p in rest)
        if not (isinstance(head, Sel) and isinstance(head.a, Ref) and head.a.var is v and head.lsb == 0
                and head.width == v.width - k and 0 < k <= 64):
            return None
        if any(is_wide(p.width) for p in rest):
            return None
        low = rest[0] if len(rest) == 1 else Concat(rest)
        return k, self.ex(low)

    def _write_through(self, v: Var):
        """In per-mtask kernels, store a stored variable as soon as it is
        written: its register can then be released instead of staying live
        until the kernel's end (thousands of values in a large mtask)."""
        if self.write_through and v in self.comp.stored:
            self.store(v)
            self.dirty.discard(v)

    def mem_assign(self, s: Assign):
        lhs = s.lhs
        m = lhs.var
        ix = self.idx_tmp(lhs.idx)
        n = m.words
        # value of the whole element after the write
        if lhs.lsb is None:
            if is_wide(m.width):
                val = self.exw(s.rhs)
                vals = [f"{val}[{k}]" for k in range(n)]
            else:
                vals = [self.ex(s.rhs)]
        else:
            # partial element write: read-modify-write of the element
            lsb = str(lhs.lsb) if isinstance(lhs.lsb, int) else self.amount(lhs.lsb)
            if is_wide(m.width):
                t = self.tmp("w")
                self.emit(f"uint32_t {t}[{n}];")
                self.emit(f"if ({ix} < {m.depth}u) {{")
                for k in range(n):
                    self.emit(f"    {t}[{k}] = {self.mem_elem(m, ix, k)};")
                self.emit(f"}} else {{ rfw_zero({t}, {m.width}); }}")
                if is_wide(lhs.width):
                    self.emit(f"rfw_putw({t}, {m.width}, {lsb}, {self.exw(s.rhs)}, {lhs.width});")
                else:
                    self.emit(f"rfw_put64({t}, {m.width}, {lsb}, {lhs.width}, (uint64_t)({self.ex(s.rhs)}));")
                vals = [f"{t}[{k}]" for k in range(n)]
            else:
                old = f"({ix} < {m.depth}u ? (uint64_t){self.mem_elem(m, ix)} : 0ull)"
                vals = [f"rf_ins64({old}, {m.width}, {lsb}, {lhs.width}, (uint64_t)({self.ex(s.rhs)}))"]
        c = class_of(m)
        if s.nb:
            # deferred to the end of the clocked block
            k = len(self.mem_nb)
            f = f"mwf{self.uid}_{k}"
            self.mem_nb.append((f, m, n))
            self.emit_top(f"uint32_t {f} = 0; uint64_t {f}_i = 0; uint32_t {f}_v[{max(n, 2)}];")
            self.emit(f"{f} = 1u; {f}_i = {ix};")
            if n == 1 and class_of(m) == "Q":
                self.emit(f"{f}_v[0] = (uint32_t)({vals[0]}); {f}_v[1] = (uint32_t)((uint64_t)({vals[0]}) >> 32);")
            else:
                for k, x in enumerate(vals):
                    self.emit(f"{f}_v[{k}] = (uint32_t)({x});")
        else:
            self.emit(f"if ({ix} < {m.depth}u) {{")
            for k, x in enumerate(vals):
                self.emit(f"    {self.mem_elem(m, ix, k)} = ({CTY[c]})({x});")
            self.emit("}")

    def commit_mem_nb(self):
        for f, m, n in self.mem_nb:
            c = class_of(m)
            self.emit(f"if ({f} && {f}_i < {m.depth}u) {{")
            if n == 1 and c == "Q":
                self.emit(f"    {self.mem_elem(m, f + '_i')} = (uint64_t){f}_v[0] | ((uint64_t){f}_v[1] << 32);")
            else:
                for k in range(n):
                    self.emit(f"    {self.mem_elem(m, f + '_i', k)} = ({CTY[c]}){f}_v[{k}];")
            self.emit("}")
            self.emit(f"{f} = 0u;")
        self.mem_nb = []

    def stmts(self, ss: List[Stmt]):
        for s in ss:
            if isinstance(s, Assign):
                self.assign(s)
            elif isinstance(s, If):
                c = self.ex(s.cond)
                saved = set(self.avail)
                self.emit(f"if ({c}) {{")
                self.ind += 1
                self.stmts(s.then)
                self.ind -= 1
                a1 = self.avail
                self.avail = set(saved)
                if s.els:
                    self.emit("} else {")
                    self.ind += 1
                    self.stmts(s.els)
                    self.ind -= 1
                self.emit("}")
                self.avail = a1 & self.avail
            elif isinstance(s, Case):
                self.case(s)
            else:
                raise CompileError(f"statement {s!r} not supported")

    def case(self, s: Case):
        w = s.expr.width
        if is_wide(w):
            sel = self.exw(s.expr)
        else:
            t = self.tmp("cs")
            self.emit(f"const {cty(w)} {t} = {self.ex(s.expr)};")
            sel = t
        saved = set(self.avail)
        outs = []
        first = True
        default = None
        for labels, body in s.items:
            if labels is None:
                if default is None:
                    default = body
                continue
            conds = []
            for val, care in labels:
                if is_wide(w):
                    if care != mask(w):
                        raise CompileError("casez on wide selectors is not supported")
                    ws = ", ".join(f"0x{(val >> (32 * i)) & 0xffffffff:x}u" for i in range(words_of(w)))
                    k = self.tmp("cl")
                    self.emit(f"const uint32_t {k}[{words_of(w)}] = {{{ws}}};")
                    conds.append(f"rfw_eq({sel}, {k}, {w})")
                elif care == mask(w):
                    conds.append(f"{sel} == {lit(val, w)}")
                else:
                    conds.append(f"(({sel} ^ {lit(val, w)}) & {lit(care, w)}) == 0")
            kw = "if" if first else "} else if"
            self.emit(f"{kw} ({' || '.join(conds)}) {{")
            first = False
            self.avail = set(saved)
            self.ind += 1
            self.stmts(body)
            self.ind -= 1
            outs.append(self.avail)
        if default is not None:
            if first:
                self.avail = set(saved)
                self.stmts(default)
                outs.append(self.avail)
            else:
                self.emit("} else {")
                self.avail = set(saved)
                self.ind += 1
                self.stmts(default)
                self.ind -= 1
                outs.append(self.avail)
                self.emit("}")
        else:
            if not first:
                self.emit("}")
            outs.append(set(saved))
        res = set(saved)
        for o in outs:
            res &= o
        self.avail = res

    # ------------------------------------------------------- processes
    def seq_group(self, procs, accs, uid: str):
        """Emit a group of clocked processes that share one edge event: every
        statement reads the values from before the edge, so non-blocking
        writes (to variables and memories) are committed only after the last
        process of the group."""
        self.uid = uid
        need_old = set()
        nb_writes = set()
        for acc in accs:
            need_old |= acc.need_old
            nb_writes |= acc.nb_writes
        # old values of conditionally / partially written targets
        for v in sorted(need_old, key=lambda x: x.id):
            if not v.is_mem:
                self.ensure(v)
        # a non-blocking target needs a temporary only if the group reads it
        nbv = [v for v in sorted(nb_writes, key=lambda x: x.id)
               if not v.is_mem and (self.nb_needed is None or v in self.nb_needed)]
        for v in nbv:
            self.declare(v)
            t = f"nb{v.id}_{uid}"
            if is_wide(v.width):
                self.emit(f"uint32_t {t}[{v.words}];")
                if v in need_old:
                    self.emit(f"rfw_copy({t}, {self.name(v)}, {v.width});")
            else:
                if v in need_old:
                    self.emit(f"{cty(v.width)} {t} = {self.get(v)};")
                else:
                    self.emit(f"{cty(v.width)} {t} = 0;")
            self.nb[v] = t
        for k, p in enumerate(procs):
            self.uid = f"{uid}_{k}"
            self.stmts(p.stmts)
        for v in nbv:
            t = self.nb.pop(v)
            if is_wide(v.width):
                self.emit(f"rfw_copy({self.name(v)}, {t}, {v.width});")
            else:
                self.emit(self.set(v, t))
            self.dirty.add(v)
            self._write_through(v)
        self.commit_mem_nb()

    def process(self, p: Process, acc, uid: str):
        """Emit one process body (blocking/NBA semantics)."""
        self.uid = uid
        if p.kind == "seq":
            self.seq_group([p], [acc], uid)
        else:
            # comb: vars whose old value is needed must be loaded first
            for v in sorted(acc.need_old, key=lambda x: x.id):
                if not v.is_mem:
                    self.ensure(v)
            self.stmts(p.stmts)
            for v in acc.writes:
                if not v.is_mem:
                    self.avail.add(v)


def edge_cond(gen: Gen, edges) -> str:
    conds = []
    for kind, clk in edges:
        last = gen.comp.clklast[clk]
        c = gen.ex(Ref(clk))
        l = gen.ex(Ref(last))
        if kind == "posedge":
            conds.append(f"(({c} & ~{l}) & 1u)")
        else:
            conds.append(f"((~{c} & {l}) & 1u)")
    return " | ".join(conds)
