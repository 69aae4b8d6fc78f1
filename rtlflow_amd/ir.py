"""Design IR: a flattened, width-exact word-level netlist.

This is the structure the reference keeps in Verilator's AST after V3Width /
V3Const / V3Gate (src/V3AstNodes.h): variables with a bit width and a storage
class, and processes (continuous assigns, combinational blocks, clocked blocks,
initial blocks) whose statements hold width-exact expression trees.

Value convention: every expression node produces an unsigned bit pattern of
exactly ``width`` bits ("clean" in Verilator's terms); signedness is carried by
the operator (``lts``, ``sra``, ``sext`` ...), never by the value.

Storage classes follow the reference's SoA allocator split
(src/V3EmitC.cpp:4597-4616): width <= 8 -> C (u8), <= 16 -> S (u16),
<= 32 -> I (u32), <= 64 -> Q (u64), wider -> W (32-bit words).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple, Union


def mask(w: int) -> int:
    return (1 << w) - 1


def words_of(w: int) -> int:
    return (w + 31) // 32


def storage_class(width: int) -> str:
    if width <= 8:
        return "C"
    if width <= 16:
        return "S"
    if width <= 32:
        return "I"
    if width <= 64:
        return "Q"
    return "W"


CLASS_BYTES = {"C": 1, "S": 2, "I": 4, "Q": 8, "W": 4}


class Var:
    """A design variable. ``depth`` is set for memories (unpacked arrays)."""

    __slots__ = ("name", "width", "signed", "kind", "depth", "lo", "id", "init",
                 "is_port", "public")

    def __init__(self, name: str, width: int, signed: bool = False, kind: str = "wire",
                 depth: Optional[int] = None, lo: int = 0):
        self.name = name
        self.width = width
        self.signed = signed
        self.kind = kind          # input / output / reg / wire
        self.depth = depth
        self.lo = lo              # lowest unpacked index (memory index offset)
        self.id = -1
        self.init = None          # constant initial value (scalars) or None
        self.is_port = kind in ("input", "output")
        self.public = False

    @property
    def cls(self) -> str:
        return storage_class(self.width)

    @property
    def words(self) -> int:
        """Storage slots per element in the class arrays."""
        return words_of(self.width) if self.width > 64 else 1

    @property
    def is_mem(self) -> bool:
        return self.depth is not None

    @property
    def slots(self) -> int:
        return self.words * (self.depth if self.depth is not None else 1)

    def __repr__(self):
        d = f"[{self.depth}]" if self.depth is not None else ""
        return f"Var({self.name}:{self.width}{d})"


# ----------------------------------------------------------------------------
# Expressions
# ----------------------------------------------------------------------------

class Expr:
    __slots__ = ("width",)

    def children(self) -> Sequence["Expr"]:
        return ()


class Const(Expr):
    __slots__ = ("value",)

    def __init__(self, width: int, value: int):
        self.width = width
        self.value = value & mask(width)

    def __repr__(self):
        return f"{self.width}'h{self.value:x}"


class Ref(Expr):
    __slots__ = ("var",)

    def __init__(self, var: Var):
        assert not var.is_mem
        self.width = var.width
        self.var = var

    def __repr__(self):
        return self.var.name


class MemRd(Expr):
    """Memory element read; out-of-range index reads 0 (two-state X)."""
    __slots__ = ("var", "idx")

    def __init__(self, var: Var, idx: Expr):
        self.width = var.width
        self.var = var
        self.idx = idx

    def children(self):
        return (self.idx,)

    def __repr__(self):
        return f"{self.var.name}[{self.idx}]"


class Sel(Expr):
    """Bits [lsb +: width] of ``a``; bits at or above a.width read 0.

    ``lsb`` is an int (constant select) or an Expr (dynamic select)."""
    __slots__ = ("a", "lsb")

    def __init__(self, a: Expr, lsb: Union[int, Expr], width: int):
        self.width = width
        self.a = a
        self.lsb = lsb

    def children(self):
        return (self.a,) if isinstance(self.lsb, int) else (self.a, self.lsb)

    def __repr__(self):
        return f"{self.a}[{self.lsb}+:{self.width}]"


class Concat(Expr):
    """Concatenation, most significant part first."""
    __slots__ = ("parts",)

    def __init__(self, parts: List[Expr]):
        self.parts = parts
        self.width = sum(p.width for p in parts)

    def children(self):
        return self.parts

    def __repr__(self):
        return "{" + ", ".join(map(repr, self.parts)) + "}"


UNARY_OPS = {"not", "neg", "redand", "redor", "redxor", "lognot", "zext", "sext"}
BINARY_SAME = {"and", "or", "xor", "xnor", "add", "sub", "mul", "divu", "modu", "divs", "mods"}
COMPARE_OPS = {"eq", "ne", "ltu", "leu", "gtu", "geu", "lts", "les", "gts", "ges"}
LOGIC_OPS = {"logand", "logor"}
SHIFT_OPS = {"shl", "shr", "sra"}


class Op(Expr):
    """Operator node.

    * same-width binary ops: result width == operand widths
    * compares / logand / logor / reductions / lognot: 1-bit result
    * shifts: result width == width of args[0]; args[1] is any width
    * zext / sext: args[0] extended to ``width``
    """
    __slots__ = ("op", "args")

    def __init__(self, op: str, args: List[Expr], width: int):
        self.op = op
        self.args = args
        self.width = width

    def children(self):
        return self.args

    def __repr__(self):
        return f"{self.op}({', '.join(map(repr, self.args))})"


class Cond(Expr):
    __slots__ = ("c", "a", "b")

    def __init__(self, c: Expr, a: Expr, b: Expr):
        assert a.width == b.width, (a, b)
        assert c.width == 1
        self.width = a.width
        self.c = c
        self.a = a
        self.b = b

    def children(self):
        return (self.c, self.a, self.b)

    def __repr__(self):
        return f"({self.c} ? {self.a} : {self.b})"


# ----------------------------------------------------------------------------
# Statements
# ----------------------------------------------------------------------------

class Lhs:
    """Assignment target: whole var, part of a var, or (part of) a memory element."""
    __slots__ = ("var", "idx", "lsb", "width")

    def __init__(self, var: Var, idx: Optional[Expr] = None,
                 lsb: Union[None, int, Expr] = None, width: Optional[int] = None):
        self.var = var
        self.idx = idx          # memory element index (memories only)
        self.lsb = lsb          # None = whole element
        self.width = var.width if lsb is None else width

    @property
    def whole(self) -> bool:
        return self.lsb is None and self.idx is None

    def exprs(self) -> List[Expr]:
        out = []
        if self.idx is not None:
            out.append(self.idx)
        if isinstance(self.lsb, Expr):
            out.append(self.lsb)
        return out

    def __repr__(self):
        s = self.var.name
        if self.idx is not None:
            s += f"[{self.idx}]"
        if self.lsb is not None:
            s += f"[{self.lsb}+:{self.width}]"
        return s


class Stmt:
    pass


class Assign(Stmt):
    __slots__ = ("lhs", "rhs", "nb")

    def __init__(self, lhs: Lhs, rhs: Expr, nb: bool):
        assert lhs.width == rhs.width, (lhs, rhs, lhs.width, rhs.width)
        self.lhs = lhs
        self.rhs = rhs
        self.nb = nb

    def __repr__(self):
        return f"{self.lhs} {'<=' if self.nb else '='} {self.rhs}"


class If(Stmt):
    __slots__ = ("cond", "then", "els")

    def __init__(self, cond: Expr, then: List[Stmt], els: List[Stmt]):
        assert cond.width == 1
        self.cond = cond
        self.then = then
        self.els = els


class Case(Stmt):
    """``items``: list of (labels, stmts); labels is a list of (value, care_mask)
    pairs (care_mask has 1s where bits are compared, for casez/casex), or None
    for the default item. First matching item wins."""
    __slots__ = ("expr", "items")

    def __init__(self, expr: Expr, items: List[Tuple[Optional[List[Tuple[int, int]]], List[Stmt]]]):
        self.expr = expr
        self.items = items


# ----------------------------------------------------------------------------
# Processes and design
# ----------------------------------------------------------------------------

class Process:
    """A logic vertex. kind: 'comb' (assign / always @*), 'seq' (clocked),
    'initial'. For 'seq', ``edges`` lists (edge, clock var)."""

    __slots__ = ("kind", "stmts", "edges", "id", "name", "reads", "writes", "mem_reads",
                 "mem_writes", "line")

    def __init__(self, kind: str, stmts: List[Stmt], edges=None, name: str = "", line: int = 0):
        self.kind = kind
        self.stmts = stmts
        self.edges = edges or []
        self.id = -1
        self.name = name
        self.line = line
        self.reads: set = set()
        self.writes: set = set()

    def __repr__(self):
        return f"Process#{self.id}({self.kind} {self.name})"


@dataclass
class Design:
    name: str
    vars: List[Var]
    processes: List[Process]
    inputs: List[Var]
    outputs: List[Var]
    clocks: List[Var] = field(default_factory=list)

    def var(self, name: str) -> Var:
        for v in self.vars:
            if v.name == name:
                return v
        raise KeyError(name)

    def finalize(self):
        for i, v in enumerate(self.vars):
            v.id = i
        for i, p in enumerate(self.processes):
            p.id = i
            p.reads, p.writes = process_access(p)
        clocks = []
        for p in self.processes:
            for _, clk in p.edges:
                if clk not in clocks:
                    clocks.append(clk)
        self.clocks = clocks


# ----------------------------------------------------------------------------
# Traversal helpers
# ----------------------------------------------------------------------------

def walk_expr(e: Expr):
    stack = [e]
    while stack:
        x = stack.pop()
        yield x
        stack.extend(x.children())


def expr_reads(e: Expr, out: set):
    for x in walk_expr(e):
        if isinstance(x, Ref):
            out.add(x.var)
        elif isinstance(x, MemRd):
            out.add(x.var)


def walk_stmts(stmts: List[Stmt]):
    for s in stmts:
        yield s
        if isinstance(s, If):
            yield from walk_stmts(s.then)
            yield from walk_stmts(s.els)
        elif isinstance(s, Case):
            for _, body in s.items:
                yield from walk_stmts(body)


def stmt_exprs(s: Stmt) -> List[Expr]:
    if isinstance(s, Assign):
        return [s.rhs] + s.lhs.exprs()
    if isinstance(s, If):
        return [s.cond]
    if isinstance(s, Case):
        return [s.expr]
    return []


def process_access(p: Process):
    reads: set = set()
    writes: set = set()
    for s in walk_stmts(p.stmts):
        for e in stmt_exprs(s):
            expr_reads(e, reads)
        if isinstance(s, Assign):
            writes.add(s.lhs.var)
            if not s.lhs.whole:
                # partial writes keep the other bits: a read-modify-write
                reads.add(s.lhs.var)
    for _, clk in p.edges:
        reads.add(clk)
    return reads, writes


def count_nodes(e: Expr) -> int:
    return sum(1 for _ in walk_expr(e))
