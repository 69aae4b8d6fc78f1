"""Bit-plane form: a design lowered to single-bit logic, evaluated for 64
stimuli per 64-bit word.

For designs whose logic is bit-level in nature (gate netlists: AND / OR /
XOR / NAND / MUX over 1-bit signals, constant selects, concatenations), the
per-lane form spends one 8-bit slot and several instructions per gate per
stimulus. Here every signal bit becomes a *plane*: one 64-bit word whose bit
``t`` is the value for stimulus ``t`` of a 64-stimulus group (one wavefront
of lanes). One gate is then one bit-select instruction pair for 64 stimuli.

``lower(comp)`` symbolically executes the design's scheduled processes into a
hash-consed DAG of single-bit nodes (AND, XOR, MUX over complementable
literals), following the cycle protocol the per-lane kernels implement:

  A(c)     inputs of cycle c, clk = 0: combinational logic -> next state (D)
  commit   state <- D for active lanes (the clk = 1 edge: every clocked block
           reads pre-edge values, non-blocking semantics)
  B(c)     clk = 1: combinational logic -> hashed output bits
  F        after the last cycle: every stored variable, for the write-back
           to the lane-minor state arrays

B(c) and A(c + 1) both read the state after cycle c's edge, so the kernel
runs them as one merged pass (program M) over two sets of input planes.

Each phase keeps only the cone of what it produces and is levelised; the
runtime kernel (csrc/rf_bitplane.hip) evaluates one level per barrier with
every node value in LDS. Nodes take LDS slots by linear-scan allocation over
levels.

Synchronous memories stay in the lane-minor arrays (their per-lane layout is
unchanged): a clocked block's ``x <= mem[addr]`` is a read port whose data
planes the kernel writes straight into the state planes of ``x`` at the
edge, and ``if (c) mem[addr] <= d`` is a write port with guard ``c``. At the
edge each lane gathers its port addresses and data from the planes, loads
the old word for every read port, then stores for every write port whose
guard is set (non-blocking semantics).

Anything outside the supported subset (asynchronous memory reads, memory
read data used by logic before the edge, memories wider than 64 bits,
multiplies wider than 16 bits, division, combinational cycles, several
clocks, negedge blocks, latches) raises ``Unsupported`` and the design keeps
only the per-lane forms.

``encode(prog)`` packs the program into the u32 blob the design library
exports (``rf_design.bitplane``); its layout is documented on ``encode``.
"""
from __future__ import annotations

from typing import Callable, Dict, List, Optional, Sequence, Tuple

from .ir import (Assign, Case, Concat, Cond, Const, Expr, If, Lhs, MemRd, Op, Ref, Sel, Var)

Lit = int          # 2 * node + inverted
F, T = 0, 1        # constant literals (node 0 is the constant 0)

MAX_SLOTS = 20000  # 160 KB of LDS at 8 bytes per slot
MAX_MUL_WIDTH = 16


class Unsupported(Exception):
    pass


LEAVES = ("c0", "pi", "pib", "st", "mrd")


class Net:
    """Hash-consed single-bit DAG. Node kinds: 'c0', 'pi', 'st', 'and',
    'xor', 'mux'. Operands are literals."""

    def __init__(self):
        self.kind: List[str] = ["c0"]
        self.ops: List[Tuple[int, ...]] = [()]
        self.info: List[object] = [None]
        self.memo: Dict[Tuple, int] = {}

    def leaf(self, kind: str, info) -> Lit:
        n = len(self.kind)
        self.kind.append(kind)
        self.ops.append(())
        self.info.append(info)
        return 2 * n

    def _node(self, kind: str, ops: Tuple[int, ...]) -> Lit:
        key = (kind,) + ops
        n = self.memo.get(key)
        if n is None:
            n = len(self.kind)
            self.kind.append(kind)
            self.ops.append(ops)
            self.info.append(None)
            self.memo[key] = n
        return 2 * n

    # -------------------------------------------------------------- gates
    def AND(self, a: Lit, b: Lit) -> Lit:
        if a == F or b == F or a == (b ^ 1):
            return F
        if a == T:
            return b
        if b == T or a == b:
            return a
        if a > b:
            a, b = b, a
        return self._node("and", (a, b))

    def OR(self, a: Lit, b: Lit) -> Lit:
        return self.AND(a ^ 1, b ^ 1) ^ 1

    def XOR(self, a: Lit, b: Lit) -> Lit:
        inv = (a & 1) ^ (b & 1)
        a &= ~1
        b &= ~1
        if a == b:
            return inv
        if a == F:
            return b ^ inv
        if b == F:
            return a ^ inv
        if a > b:
            a, b = b, a
        return self._node("xor", (a, b)) ^ inv

    def NOT(self, a: Lit) -> Lit:
        return a ^ 1

    def MUX(self, s: Lit, a: Lit, b: Lit) -> Lit:
        """s ? a : b"""
        if s == T:
            return a
        if s == F:
            return b
        if a == b:
            return a
        if s & 1:
            s, a, b = s ^ 1, b, a
        if a == T and b == F:
            return s
        if a == F and b == T:
            return s ^ 1
        if a == F:
            return self.AND(s ^ 1, b)
        if b == F:
            return self.AND(s, a)
        if a == T:
            return self.OR(s, b)
        if b == T:
            return self.OR(s ^ 1, a)
        if a == (b ^ 1):
            return self.XOR(s, b)
        if a == s:
            return self.OR(s, b)
        if a == (s ^ 1):
            return self.AND(s ^ 1, b)
        if b == s:
            return self.AND(s, a)
        if b == (s ^ 1):
            return self.OR(s ^ 1, a)
        inv = a & 1
        return self._node("mux", (s, a ^ inv, b ^ inv)) ^ inv

    # ------------------------------------------------------ word helpers
    def const(self, width: int, value: int) -> List[Lit]:
        return [T if (value >> i) & 1 else F for i in range(width)]

    def redor(self, bits: Sequence[Lit]) -> Lit:
        return self._tree(list(bits), self.OR, F)

    def redand(self, bits: Sequence[Lit]) -> Lit:
        return self._tree(list(bits), self.AND, T)

    def redxor(self, bits: Sequence[Lit]) -> Lit:
        return self._tree(list(bits), self.XOR, F)

    def _tree(self, xs: List[Lit], f, unit: Lit) -> Lit:
        if not xs:
            return unit
        while len(xs) > 1:
            nxt = [f(xs[i], xs[i + 1]) for i in range(0, len(xs) - 1, 2)]
            if len(xs) & 1:
                nxt.append(xs[-1])
            xs = nxt
        return xs[0]

    def add(self, a: Sequence[Lit], b: Sequence[Lit], cin: Lit = F) -> List[Lit]:
        out = []
        c = cin
        for x, y in zip(a, b):
            t = self.XOR(x, y)
            out.append(self.XOR(t, c))
            c = self.MUX(t, c, x)
        return out

    def borrow(self, a: Sequence[Lit], b: Sequence[Lit]) -> Lit:
        """1 iff a < b (unsigned): the carry out of a + ~b + 1 is 0."""
        c = T
        for x, y in zip(a, b):
            t = self.XOR(x, y ^ 1)
            c = self.MUX(t, c, x)
        return c ^ 1

    def shift(self, bits: List[Lit], amt: List[Lit], left: bool, fill: Lit) -> List[Lit]:
        w = len(bits)
        x = list(bits)
        k = 0
        while k < len(amt) and (1 << k) < w:
            s = 1 << k
            if left:
                y = [fill] * s + x[:w - s]
            else:
                y = x[s:] + [fill] * s
            x = [self.MUX(amt[k], yi, xi) for xi, yi in zip(x, y)]
            k += 1
        # any higher amount bit set: everything shifted out
        hi = amt[k:]
        if hi:
            big = self.redor(hi)
            x = [self.MUX(big, fill, xi) for xi in x]
        return x


def _fit(bits: List[Lit], w: int) -> List[Lit]:
    return bits[:w] + [F] * (w - len(bits))


class Lowerer:
    def __init__(self, comp):
        self.comp = comp
        self.d = comp.design
        self.net = Net()
        self.pi: Dict[Var, List[Lit]] = {}
        self.pi_b: Dict[Var, List[Lit]] = {}
        self.state: Dict[Var, List[Lit]] = {}
        # synchronous memory ports of the clocked blocks
        self.in_seq = False
        self.guard: Lit = T                  # path condition of the statement being run
        self.rports: List[Tuple[Var, List[Lit], List[Lit]]] = []   # (mem, addr, data leaves)
        self.rport_of: Dict[Tuple[int, Tuple[Lit, ...]], int] = {}
        self.wports: List[Tuple[Var, Lit, List[Lit], List[Lit]]] = []  # (mem, guard, addr, data)

    # ----------------------------------------------------------- exprs
    def ev(self, e: Expr, env: Dict[Var, List[Lit]]) -> List[Lit]:
        n = self.net
        if isinstance(e, Const):
            return n.const(e.width, e.value)
        if isinstance(e, Ref):
            v = env.get(e.var)
            if v is None:
                raise Unsupported(f"read of {e.var.name} before any value")
            return list(v)
        if isinstance(e, MemRd):
            if not self.in_seq:
                raise Unsupported("asynchronous memory read")
            m = e.var
            self._check_mem(m)
            addr = tuple(self._addr(e.idx, env))
            key = (m.id, addr)
            if key not in self.rport_of:
                k = len(self.rports)
                self.rport_of[key] = k
                self.rports.append((m, list(addr), [n.leaf("mrd", (k, i)) for i in range(m.width)]))
            return list(self.rports[self.rport_of[key]][2])
        if isinstance(e, Sel):
            a = self.ev(e.a, env)
            if isinstance(e.lsb, int):
                return _fit(a[e.lsb:], e.width)
            amt = self.ev(e.lsb, env)
            wide = a + [F] * e.width
            return n.shift(wide, amt, left=False, fill=F)[:e.width]
        if isinstance(e, Concat):
            out: List[Lit] = []
            for p in reversed(e.parts):
                out += self.ev(p, env)
            return out
        if isinstance(e, Cond):
            c = self.ev(e.c, env)[0]
            a = self.ev(e.a, env)
            b = self.ev(e.b, env)
            return [n.MUX(c, x, y) for x, y in zip(a, b)]
        if isinstance(e, Op):
            return self.op(e, env)
        raise Unsupported(f"expression {type(e).__name__}")

    def op(self, e: Op, env) -> List[Lit]:
        n = self.net
        o = e.op
        args = [self.ev(x, env) for x in e.args]
        w = e.width
        if o == "not":
            return [x ^ 1 for x in args[0]]
        if o in ("and", "or", "xor", "xnor"):
            f = {"and": n.AND, "or": n.OR, "xor": n.XOR,
                 "xnor": lambda x, y: n.XOR(x, y) ^ 1}[o]
            return [f(x, y) for x, y in zip(args[0], args[1])]
        if o == "add":
            return n.add(args[0], args[1])
        if o == "sub":
            return n.add(args[0], [x ^ 1 for x in args[1]], T)
        if o == "neg":
            return n.add(n.const(w, 0), [x ^ 1 for x in args[0]], T)
        if o == "mul":
            if w > MAX_MUL_WIDTH:
                raise Unsupported("wide multiply")
            acc = n.const(w, 0)
            a, b = args
            for i in range(w):
                part = [F] * i + [n.AND(x, b[i]) for x in a[:w - i]]
                acc = n.add(acc, part)
            return acc
        if o in ("divu", "modu", "divs", "mods"):
            raise Unsupported("division")
        if o == "redand":
            return [n.redand(args[0])]
        if o == "redor":
            return [n.redor(args[0])]
        if o == "redxor":
            return [n.redxor(args[0])]
        if o == "lognot":
            return [n.redor(args[0]) ^ 1]
        if o == "logand":
            return [n.AND(n.redor(args[0]), n.redor(args[1]))]
        if o == "logor":
            return [n.OR(n.redor(args[0]), n.redor(args[1]))]
        if o == "zext":
            return _fit(args[0], w)
        if o == "sext":
            a = args[0]
            return a + [a[-1]] * (w - len(a))
        if o in ("eq", "ne"):
            d = n.redor([n.XOR(x, y) for x, y in zip(args[0], args[1])])
            return [d ^ 1 if o == "eq" else d]
        if o in ("ltu", "leu", "gtu", "geu", "lts", "les", "gts", "ges"):
            a, b = args
            if o[-1] == "s":
                a = a[:-1] + [a[-1] ^ 1]
                b = b[:-1] + [b[-1] ^ 1]
            base = o[:2]
            if base == "lt":
                r = n.borrow(a, b)
            elif base == "gt":
                r = n.borrow(b, a)
            elif base == "le":
                r = n.borrow(b, a) ^ 1
            else:
                r = n.borrow(a, b) ^ 1
            return [r]
        if o in ("shl", "shr", "sra"):
            a, amt = args
            fill = a[-1] if o == "sra" else F
            if isinstance(e.args[1], Const):
                k = e.args[1].value
                if k >= w:
                    return [fill] * w
                if o == "shl":
                    return [F] * k + a[:w - k]
                return a[k:] + [fill] * k
            return n.shift(a, amt, left=(o == "shl"), fill=fill)
        raise Unsupported(f"operator {o}")

    # ------------------------------------------------------- statements
    def _check_mem(self, m: Var) -> None:
        if m.width > 64:
            raise Unsupported(f"memory {m.name} wider than 64 bits")

    def _addr(self, idx: Expr, env) -> List[Lit]:
        a = self.ev(idx, env)
        if len(a) > 64:
            raise Unsupported("memory index wider than 64 bits")
        return a

    def write(self, env, lhs: Lhs, val: List[Lit], rd_env) -> None:
        v = lhs.var
        if v.is_mem or lhs.idx is not None:
            raise Unsupported("memory write")
        if lhs.whole:
            env[v] = list(val)
            return
        old = env.get(v)
        if old is None:
            raise Unsupported(f"partial write of {v.name} without a prior value")
        old = list(old)
        if isinstance(lhs.lsb, int):
            for i, b in enumerate(val):
                if lhs.lsb + i < v.width:
                    old[lhs.lsb + i] = b
            env[v] = old
            return
        n = self.net
        amt = self.ev(lhs.lsb, rd_env)
        shifted = n.shift(_fit(val, v.width), amt, left=True, fill=F)
        m = n.shift(_fit([T] * lhs.width, v.width), amt, left=True, fill=F)
        env[v] = [n.MUX(mi, si, oi) for mi, si, oi in zip(m, shifted, old)]

    def merge(self, c: Lit, a: Dict[Var, List[Lit]], b: Dict[Var, List[Lit]],
              base: Dict[Var, List[Lit]], keep_old: bool = False) -> Dict[Var, List[Lit]]:
        """Join two branches. A variable set on one branch only keeps its old
        value on the other: the pre-edge state for non-blocking writes
        (``keep_old``), otherwise a latch, which this form does not model."""
        n = self.net
        out = dict(base)
        for v in set(a) | set(b):
            x, y = a.get(v), b.get(v)
            if x is y:
                out[v] = x
                continue
            if keep_old:
                x = x if x is not None else self.state_bits(v)
                y = y if y is not None else self.state_bits(v)
            elif x is None or y is None:
                raise Unsupported(f"{v.name} assigned on one branch only (latch)")
            out[v] = [n.MUX(c, p, q) for p, q in zip(x, y)]
        return out

    def run(self, stmts, env, nba, seq: bool):
        """Execute statements symbolically. env: blocking view (reads and
        blocking writes); nba: pending non-blocking values (seq only)."""
        n = self.net
        for s in stmts:
            if isinstance(s, Assign) and s.lhs.var.is_mem:
                m = s.lhs.var
                if not (seq and s.nb):
                    raise Unsupported(f"memory {m.name} written outside a clocked non-blocking assignment")
                if s.lhs.lsb is not None:
                    raise Unsupported(f"partial write of an element of {m.name}")
                self._check_mem(m)
                addr = self._addr(s.lhs.idx, env)
                data = _fit(self.ev(s.rhs, env), m.width)
                self.wports.append((m, self.guard, addr, data))
            elif isinstance(s, Assign):
                val = self.ev(s.rhs, env)
                if s.nb:
                    if not seq:
                        raise Unsupported("non-blocking assignment in combinational logic")
                    if s.lhs.var not in nba:
                        nba[s.lhs.var] = list(self.state_bits(s.lhs.var))
                    self.write(nba, s.lhs, val, env)
                else:
                    self.write(env, s.lhs, val, env)
            elif isinstance(s, If):
                c = self.ev(s.cond, env)[0]
                self._branch(c, s.then, s.els, env, nba, seq)
            elif isinstance(s, Case):
                x = self.ev(s.expr, env)
                self._case(x, s.items, 0, env, nba, seq)
            else:
                raise Unsupported(f"statement {type(s).__name__}")

    def _branch(self, c, then, els, env, nba, seq):
        n = self.net
        g = self.guard
        e1, n1 = dict(env), dict(nba)
        self.guard = n.AND(g, c)
        self.run(then, e1, n1, seq)
        e2, n2 = dict(env), dict(nba)
        self.guard = n.AND(g, c ^ 1)
        self.run(els, e2, n2, seq)
        self.guard = g
        env.update(self.merge(c, e1, e2, env))
        nba.update(self.merge(c, n1, n2, nba, keep_old=True))

    def _case(self, x, items, k, env, nba, seq):
        n = self.net
        if k >= len(items):
            return
        labels, body = items[k]
        if labels is None:
            self.run(body, env, nba, seq)
            return
        hits = []
        for value, care in labels:
            d = [n.XOR(b, T if (value >> i) & 1 else F) for i, b in enumerate(x) if (care >> i) & 1]
            hits.append(n.redor(d) ^ 1)
        c = n.redor(hits)
        g = self.guard
        e1, n1 = dict(env), dict(nba)
        self.guard = n.AND(g, c)
        self.run(body, e1, n1, seq)
        e2, n2 = dict(env), dict(nba)
        self.guard = n.AND(g, c ^ 1)
        self._case(x, items, k + 1, e2, n2, seq)
        self.guard = g
        env.update(self.merge(c, e1, e2, env))
        nba.update(self.merge(c, n1, n2, nba, keep_old=True))

    def state_bits(self, v: Var) -> List[Lit]:
        if v not in self.state:
            raise Unsupported(f"non-blocking write to {v.name}, which is not clocked state")
        return self.state[v]

    # -------------------------------------------------------- the design
    def comb_env(self, clk_val: Lit, pi: Dict[Var, List[Lit]]) -> Dict[Var, List[Lit]]:
        comp = self.comp
        env: Dict[Var, List[Lit]] = {}
        # variables no process drives and nothing stores read as zero
        for v in self.d.vars:
            if not v.is_mem:
                env[v] = [F] * v.width
        env.update(pi)
        env.update(self.state)
        for c in self.d.clocks:
            env[c] = [clk_val]
        for vid in comp.sch.topo:
            vx = comp.sch.vertices[vid]
            if vx.kind != "comb":
                continue
            for p in vx.procs:
                local = dict(env)
                for v in vx.writes:
                    local.pop(v, None)
                self.run(p.stmts, local, {}, seq=False)
                for v in p.writes:
                    if v.is_mem:
                        raise Unsupported("memory written by combinational logic")
                    if v not in local:
                        raise Unsupported(f"{v.name} not assigned")
                    env[v] = local[v]
        return env

    def lower(self):
        comp = self.comp
        d = self.d
        sch = comp.sch
        n = self.net
        if sch.change_vars:
            raise Unsupported("combinational cycle (settle loop)")
        if len(d.clocks) > 1:
            raise Unsupported("several clocks")
        clk = d.clocks[0] if d.clocks else None
        seq_vx = [sch.vertices[vid] for vid in sch.topo if sch.vertices[vid].kind == "seq"]
        for vx in seq_vx:
            for p in vx.procs:
                if any(edge != "posedge" or c is not clk for edge, c in p.edges):
                    raise Unsupported("clocked block not on the posedge of the clock")
        # leaves
        # two sets of input leaves: the clk = 0 settle of cycle n + 1 and the
        # clk = 1 settle of cycle n run in one pass over different inputs
        for v in d.vars:
            if v.kind == "input" and v not in d.clocks:
                self.pi[v] = [n.leaf("pi", (v, i)) for i in range(v.width)]
                self.pi_b[v] = [n.leaf("pib", (v, i)) for i in range(v.width)]
        # state: every stored variable that no combinational block drives
        # (clocked registers, and registers only initial blocks set),
        # loaded from the lane-minor arrays when a launch starts
        comb_written = set()
        for vx in sch.vertices:
            if vx.kind == "comb":
                comb_written |= vx.writes
        synth = set(comp.clklast.values())
        state_vars = [v for v in d.vars if v in comp.stored and not v.is_mem and v.kind != "input"
                      and v not in comb_written and v not in synth]
        for v in state_vars:
            self.state[v] = [n.leaf("st", (v, i)) for i in range(v.width)]
        # phase A: clk = 0 settle, then the clocked blocks read pre-edge values
        envA = self.comb_env(F, self.pi)
        nxt: Dict[Var, List[Lit]] = {}
        written_by: Dict[Var, int] = {}
        for vx in seq_vx:
            for p in vx.procs:
                for v in p.writes:
                    if v in written_by and written_by[v] != p.id:
                        raise Unsupported(f"{v.name} written by several clocked blocks")
                    written_by[v] = p.id
                env = dict(envA)
                for c in d.clocks:
                    env[c] = [T]
                nba: Dict[Var, List[Lit]] = {}
                self.in_seq, self.guard = True, T
                self.run(p.stmts, env, nba, seq=True)
                self.in_seq = False
                for v in p.writes:
                    if v in nba:
                        nxt[v] = nba[v]
                    elif v in self.state:
                        # blocking write of a reg used only inside its block
                        readers = [q for q in d.processes if v in q.reads and q is not p]
                        if readers:
                            raise Unsupported(f"blocking write of shared {v.name} in a clocked block")
                        nxt[v] = env[v]
        for v in state_vars:
            if v not in nxt:
                nxt[v] = self.state[v]
        # phase B / F: clk = 1 settle on the committed state
        envB = self.comb_env(T, self.pi_b)
        self.envB = envB
        self.nxt = nxt
        self.state_vars = state_vars
        return self


# =====================================================================
# Levelisation and slot allocation
# =====================================================================

class Phase:
    def __init__(self):
        self.levels: List[List[Tuple[int, int, int, int, int, int]]] = []


class Program:
    pass


def _cone(net: Net, roots: Sequence[Lit]) -> List[int]:
    """Internal nodes reachable from roots, in topological order."""
    seen = set()
    order = []
    stack = [(r >> 1, False) for r in roots]
    while stack:
        x, done = stack.pop()
        if done:
            order.append(x)
            continue
        if x in seen:
            continue
        seen.add(x)
        if net.kind[x] in LEAVES:
            continue
        stack.append((x, True))
        for o in net.ops[x]:
            if (o >> 1) not in seen:
                stack.append((o >> 1, False))
    return order


def build_phase(net: Net, roots: Sequence[Lit], fixed: Dict[int, int], nfixed: int,
                keep: Sequence[Lit] = (), inplace: Sequence[Tuple[int, int]] = ()
                ) -> Tuple[Phase, Dict[int, int], int]:
    """Levelise the cone of ``roots`` and give each node an LDS slot.
    ``fixed``: node -> slot for leaves. Slots of ``keep`` nodes stay live to
    the end of the phase. ``inplace``: (next-state node, state leaf) pairs; a
    next-state node whose state leaf has no reader at or after its level
    writes the state slot directly (the pair is added to ``ph.inplace``).
    Returns (phase, node->slot, peak slots)."""
    order = _cone(net, roots)
    level: Dict[int, int] = {}
    for x in order:
        level[x] = 1 + max((level.get(o >> 1, 0) for o in net.ops[x]), default=0)
    nlev = max(level.values(), default=0)
    direct: Dict[int, int] = {}
    done_pairs = set()
    if inplace:
        last_read: Dict[int, int] = {}
        for x in order:
            for o in net.ops[x]:
                y = o >> 1
                if y not in level:
                    last_read[y] = max(last_read.get(y, 0), level[x])
        for d, q in inplace:
            if d in level and d not in direct and last_read.get(q, 0) < level[d]:
                direct[d] = fixed[q]
                done_pairs.add((d, q))
    last: Dict[int, int] = {}
    for x in order:
        for o in net.ops[x]:
            y = o >> 1
            if y in level:
                last[y] = max(last.get(y, 0), level[x])
    END = nlev + 1
    for r in list(roots) + list(keep):
        if (r >> 1) in level:
            last[r >> 1] = END
    by_level: List[List[int]] = [[] for _ in range(nlev + 1)]
    for x in order:
        by_level[level[x]].append(x)
    # free slots by bank residue: a run of BANK_GROUP gates takes
    # destinations of distinct residues (conflict-free writes)
    free: List[List[int]] = [[] for _ in range(BANK_MOD)]
    top = nfixed
    slot: Dict[int, int] = dict(fixed)
    release: Dict[int, List[int]] = {}
    for x, l in last.items():
        release.setdefault(l, []).append(x)
    peak = nfixed
    ph = Phase()
    ph.inplace = done_pairs

    def take(avoid: set) -> int:
        nonlocal top
        best = -1
        for r in range(BANK_MOD):
            if r not in avoid and free[r] and (best < 0 or len(free[r]) > len(free[best])):
                best = r
        if best >= 0:
            return free[best].pop()
        while top % BANK_MOD in avoid:
            free[top % BANK_MOD].append(top)
            top += 1
        top += 1
        return top - 1

    for L in range(1, nlev + 1):
        pending = []
        for x in by_level[L]:
            slot[x] = 0            # destination chosen below
            pending.append((x, _gate(net, x, slot)))
        ordered = bank_order(pending, direct)
        gates = []
        for r0 in range(0, len(ordered), BANK_GROUP):
            run = ordered[r0:r0 + BANK_GROUP]
            # the run's in-place destinations are fixed: the others avoid them
            avoid: set = {direct[x] % BANK_MOD for x, _ in run if x in direct}
            for x, g in run:
                if x in direct:
                    d = direct[x]
                else:
                    d = take(avoid)
                avoid.add(d % BANK_MOD)
                slot[x] = d
                gates.append((d,) + g[1:])
        peak = max(peak, top)
        ph.levels.append(gates)
        for y in release.get(L, []):
            if y in slot and y not in fixed and y not in direct:
                free[slot[y] % BANK_MOD].append(slot[y])
    return ph, slot, peak


# LDS bank model for the gate order of a level: 8-byte slots, slot i in the
# bank pair (2i, 2i + 1) of gfx950's 64 four-byte banks, BANK_GROUP lanes
# served together by a 64-bit access, so a run of BANK_GROUP gates is
# conflict-free when its slots differ modulo BANK_MOD
BANK_GROUP = 32
BANK_MOD = 32


def bank_runs(items: List, keys: Callable[[object], Tuple[Optional[int], ...]]) -> List:
    """Order ``items`` so that each run of BANK_GROUP consecutive items (the
    lanes one LDS access serves) spreads every access over the banks:
    ``keys(item)`` gives the item's slot per access (None: no constraint); a
    run takes an item only while, per access, the slot's residue is held by
    fewer than ``cap`` other slots of the run (equal slots are a broadcast
    and free); ``cap`` starts at 1 and rises only when no remaining item
    fits. Greedy."""
    rest = list(items)
    out = []
    while rest:
        used: Dict[int, Dict[int, set]] = {}
        run: List[int] = []
        taken = set()
        cap = 1
        while len(run) < min(BANK_GROUP, len(rest)):
            hit = -1
            for k, it in enumerate(rest):
                if k in taken:
                    continue
                ok = True
                for jj, o in enumerate(keys(it)):
                    if o is None:
                        continue
                    held = used.get(jj, {}).get(o % BANK_MOD)
                    if held and o not in held and len(held) >= cap:
                        ok = False
                        break
                if ok:
                    hit = k
                    break
            if hit < 0:
                cap += 1
                continue
            for jj, o in enumerate(keys(rest[hit])):
                if o is not None:
                    used.setdefault(jj, {}).setdefault(o % BANK_MOD, set()).add(o)
            run.append(hit)
            taken.add(hit)
        out += [rest[k] for k in sorted(run)]
        rest = [p for k, p in enumerate(rest) if k not in taken]
    return out


def bank_order(pending: List[Tuple[int, Tuple[int, int, int, int, int, int]]],
               fixed_dst: Optional[Dict[int, int]] = None
               ) -> List[Tuple[int, Tuple[int, int, int, int, int, int]]]:
    """Order one level's (node, gate) pairs by ``bank_runs`` over the operand
    reads (s, x, y) and, for gates whose destination is fixed (``fixed_dst``:
    next-state values written in place), the write. The other destinations
    are chosen afterwards, distinct per run. Level by level."""
    fd = fixed_dst or {}
    rest = sorted(pending, key=lambda p: (p[1][1], p[1][2]))
    return bank_runs(rest, lambda p: (p[1][1], p[1][2], p[1][4], fd.get(p[0])))


def _lit_slot(lit: Lit, slot: Dict[int, int]) -> Tuple[int, int]:
    return slot[lit >> 1], lit & 1


def _gate(net: Net, x: int, slot: Dict[int, int]) -> Tuple[int, int, int, int, int, int]:
    """(dst, s, xs, ix, ys, iy): dst = s ? (xs ^ ix) : (ys ^ iy)."""
    k = net.kind[x]
    ops = net.ops[x]
    if k == "and":
        a, b = ops
        s, inv = _lit_slot(a, slot)
        bs, bi = _lit_slot(b, slot)
        if inv:      # ~a ? b : 0  ==  a ? 0 : b
            return (slot[x], s, 0, 0, bs, bi)
        return (slot[x], s, bs, bi, 0, 0)
    if k == "xor":
        a, b = ops
        s, _ = _lit_slot(a, slot)
        bs, bi = _lit_slot(b, slot)
        return (slot[x], s, bs, bi ^ 1, bs, bi)
    if k == "copy":
        src, inv = _lit_slot(ops[0], slot)
        return (slot[x], src, 0, inv ^ 1, 0, inv)
    if k == "mux":
        sl, a, b = ops
        s, _ = _lit_slot(sl, slot)
        xs, xi = _lit_slot(a, slot)
        ys, yi = _lit_slot(b, slot)
        return (slot[x], s, xs, xi, ys, yi)
    raise AssertionError(k)


CLS_ID = {"C": 0, "S": 1, "I": 2, "Q": 3, "W": 2}
CLS_SIZE = {"C": 1, "S": 2, "I": 4, "Q": 8, "W": 4}


def compile_program(comp) -> Program:
    """Lower and schedule; raises Unsupported for designs outside the subset.

    Programs: A0 (clk = 0 settle of the first cycle -> next state), M (the
    clk = 1 settle of cycle n -> hashed outputs, merged with the clk = 0
    settle of cycle n + 1 -> next state; both read the state after cycle n's
    edge), B (the clk = 1 settle of the last cycle), and the write-back
    chunks F."""
    from .layout import class_of
    lw = Lowerer(comp).lower()
    net = lw.net
    d = comp.design
    lay = comp.layout
    # fixed slots: 0 = zeros, then the inputs of the clk = 0 settle (set A)
    # and of the clk = 1 settle (set B), then the state bits
    fixed: Dict[int, int] = {0: 0}
    top = 1
    in_chunks = []
    for port in comp.port_table():
        if port["dir"] != 0:
            continue
        v = port["var"]
        for j in range((v.width + 63) // 64):
            nb = min(64, v.width - 64 * j)
            sa, sb = top, top + nb
            for i in range(nb):
                fixed[lw.pi[v][64 * j + i] >> 1] = sa + i
                fixed[lw.pi_b[v][64 * j + i] >> 1] = sb + i
            in_chunks.append(((port["stim_index"] << 3) | j, nb, sa, sb))
            top += 2 * nb
    state_load = []
    for v in lw.state_vars:
        bits = lw.state[v]
        state_load.append((CLS_ID[class_of(v)], lay.loc[v], v.slots, v.width, top))
        for i, b in enumerate(bits):
            fixed[b >> 1] = top + i
        top += v.width
    nfixed = top
    copies: Dict[Lit, Lit] = {}

    def own(lit: Lit) -> Lit:
        """A node of its own holding the literal's value (not hash-consed)."""
        if lit not in copies:
            n = len(net.kind)
            net.kind.append("copy")
            net.ops.append((lit,))
            net.info.append(None)
            copies[lit] = 2 * n
        return copies[lit]

    # next-state roots: a literal that is a leaf or inverted gets a copy node
    # so the commit reads a slot of its own. A state bit whose next value is
    # a memory read port's data bit takes it from the port at the edge.
    commit_lits = []
    rtargets: List[List[Tuple[int, int]]] = [[] for _ in lw.rports]
    for v in lw.state_vars:
        for i, lit in enumerate(lw.nxt[v]):
            if net.kind[lit >> 1] == "mrd":
                port, bit = net.info[lit >> 1]
                if lit & 1:
                    raise Unsupported("inverted memory read data")
                rtargets[port].append((bit, fixed[lw.state[v][i] >> 1]))
                continue
            commit_lits.append((lw.state[v][i] >> 1, lit))
    droots = [own(lit) if (lit & 1) or net.kind[lit >> 1] in LEAVES else lit for _, lit in commit_lits]

    # memory port roots: the kernel gathers them at the edge, while the
    # commit rewrites state planes and the next inputs are generated, so an
    # input or state bit is copied
    def port_lit(lit: Lit) -> Lit:
        k = net.kind[lit >> 1]
        if k == "mrd":
            raise Unsupported("memory read data used by logic before the edge")
        return own(lit & ~1) ^ (lit & 1) if k in ("pi", "pib", "st") else lit

    mem_list: List[Var] = []
    rp_lits = [[port_lit(b) for b in addr] for _, addr, _ in lw.rports]
    wp_lits = [(port_lit(g), [port_lit(b) for b in addr], [port_lit(b) for b in data])
               for _, g, addr, data in lw.wports]
    for k, (m, _, _) in enumerate(lw.rports):
        if rtargets[k] and m not in mem_list:
            mem_list.append(m)
    for m, _, _, _ in lw.wports:
        if m not in mem_list:
            mem_list.append(m)
    proots = [b for k, lits in enumerate(rp_lits) if rtargets[k] for b in lits]
    proots += [b for g, addr, data in wp_lits for b in [g] + addr + data]
    # hashed output roots: the hash of cycle n runs after the next commit and
    # while the inputs of later cycles are written, so an output bit that is
    # an input or state bit is copied
    out_ports = [p for p in comp.port_table() if p["hashed"]]
    hbits = {}
    for p in out_ports:
        v = p["var"]
        hbits[v] = [own(b & ~1) ^ (b & 1) if net.kind[b >> 1] in ("pi", "pib", "st") else b
                    for b in lw.envB[v]]
    hroots = [b for p in out_ports for b in hbits[p["var"]]]

    def commit_of(ph, slot):
        return [(fixed[st_node], slot[r >> 1]) for (st_node, _), r in zip(commit_lits, droots)
                if (r >> 1, st_node) not in ph.inplace]

    pairs = [(r >> 1, st_node) for (st_node, _), r in zip(commit_lits, droots)]

    def hash_of(slot):
        out = []
        for p in out_ports:
            v = p["var"]
            nbytes = CLS_SIZE[class_of(v)] * (v.slots if v.width > 64 else 1)
            out.append((nbytes, v.width, [_lit_ref(b, slot) for b in hbits[v]]))
        return out

    for x in _cone(net, droots + hroots + proots):
        if any(net.kind[o >> 1] == "mrd" for o in net.ops[x]):
            raise Unsupported("memory read data used by logic before the edge")

    def ports_of(slot):
        """Per memory (in mem_list order): its read ports, then its write
        ports in statement order."""
        recs = []
        for mi, m in enumerate(mem_list):
            for k, (mk, _, _) in enumerate(lw.rports):
                if mk is m and rtargets[k]:
                    recs.append((0, mi, [_lit_ref(b, slot) for b in rp_lits[k]], sorted(rtargets[k])))
            for (mk, _, _, _), (g, addr, data) in zip(lw.wports, wp_lits):
                if mk is m:
                    recs.append((1, mi, [_lit_ref(b, slot) for b in addr], _lit_ref(g, slot),
                                 [_lit_ref(b, slot) for b in data]))
        return recs

    design_mems = [v for v in d.vars if v.is_mem and lay.has(v)]
    mem_table = []
    for m in mem_list:
        lm = m in lay.lane_major
        mem_table.append((design_mems.index(m), CLS_ID[class_of(m)], 1 if lm else 0,
                          lay.lane_major[m] if lm else lay.loc[m], m.depth * CLS_SIZE[m.cls],
                          m.depth, m.width))

    phA, slotA, peakA = build_phase(net, droots + proots, fixed, nfixed, inplace=pairs)
    phM, slotM, peakM = build_phase(net, droots + hroots + proots, fixed, nfixed, inplace=pairs)
    phB, slotB, peakB = build_phase(net, hroots, fixed, nfixed)
    # phase F: every stored variable after the last cycle, in chunks whose
    # values all stay live until their write-back; chunks are sized so that
    # phase F needs no more slots than the per-cycle phases
    wb = []
    for v in d.vars:
        if v.is_mem or not lay.has(v):
            continue
        bits = _final_bits(comp, lw, v)
        if bits is not None:
            wb.append((v, bits))
    target = max(peakA, peakM, peakB)
    k = max(1, sum(len(b) for _, b in wb))
    while True:
        groups, cur, n = [], [], 0
        for item in wb:
            if cur and n + len(item[1]) > k:
                groups.append(cur)
                cur, n = [], 0
            cur.append(item)
            n += len(item[1])
        if cur:
            groups.append(cur)
        fchunks = []
        peakF = 0
        for g in groups:
            ph, sl, pk = build_phase(net, [b for _, bits in g for b in bits], fixed, nfixed)
            entries = [(CLS_ID[class_of(v)], lay.loc[v], v.slots, v.width, [_lit_ref(b, sl) for b in bits])
                       for v, bits in g]
            # single-bit entries first: fixed-size records the kernel indexes directly
            entries.sort(key=lambda e: e[3] != 1)
            fchunks.append((ph, entries))
            peakF = max(peakF, pk)
        if peakF <= target or k <= 64:
            break
        k //= 2
    state_load.sort(key=lambda e: e[3] != 1)
    prog = Program()
    prog.nslots = max(target, peakF)
    if prog.nslots > MAX_SLOTS:
        raise Unsupported(f"{prog.nslots} live bit slots exceed LDS")
    prog.in_chunks = in_chunks
    prog.state_load = state_load
    prog.phases = [phA, phM, phB]
    # commit pairs run in parallel (a pair's next-state slot is its own, never
    # a state slot): ordered so each run of BANK_GROUP spreads its reads and
    # writes over the banks
    prog.commits = [bank_runs(commit_of(ph, sl), lambda c: c) for ph, sl in ((phA, slotA), (phM, slotM))]
    prog.hashes = [hash_of(slotM), hash_of(slotB)]
    prog.mems = mem_table
    prog.ports = [ports_of(slotA), ports_of(slotM)]
    prog.fchunks = fchunks
    prog.nodes = len(net.kind)
    prog.gates = [sum(len(l) for l in ph.levels) for ph in prog.phases] + \
        [sum(len(l) for ph, _ in fchunks for l in ph.levels)]
    prog.nlevels = [len(ph.levels) for ph in prog.phases] + [sum(len(ph.levels) for ph, _ in fchunks)]
    return prog


def _lit_ref(lit: Lit, slot: Dict[int, int]) -> int:
    """Slot of a literal with the inversion in bit 31."""
    return slot[lit >> 1] | ((lit & 1) << 31)


def _final_bits(comp, lw: Lowerer, v: Var) -> Optional[List[Lit]]:
    d = comp.design
    if v in d.clocks or v in comp.clklast.values():
        return [T]
    if v in lw.state:
        return lw.state[v]
    if v in lw.envB:
        return lw.envB[v]
    return None


# =====================================================================
# Encoding
# =====================================================================

BP_MAGIC = 0x50424652   # "RFBP"
BP_VERSION = 8
# gate records hold LDS byte offsets (slot * 8) when every slot fits 16 bits
BYTE_FORM_MAX_SLOTS = 8192
HDR_WORDS = 18
FCHUNK_HDR = 6


def gate_word(g: Tuple[int, int, int, int, int, int], byte_form: bool = False) -> int:
    dst, s, xs, ix, ys, iy = g
    if byte_form:
        return (8 * dst) | ((8 * s) << 16) | ((8 * xs | ix) << 32) | ((8 * ys | iy) << 48)
    return dst | (s << 16) | (xs << 32) | (ix << 47) | (ys << 48) | (iy << 63)


def _wb_words(entries) -> List[int]:
    out: List[int] = []
    for cls, loc, words, width, refs in entries:
        out += [cls, loc & 0xffffffff, loc >> 32, words, width] + refs
    return out


def encode(prog: Program) -> List[int]:
    """u32 words of the program blob.

    header (18 words): magic, version | byte_form << 16, nslots, n_in,
    n_state, n_commit[A0],
    n_hash, n_fchunks, nlev[A0, M, B], ngates[A0, M, B], n_state1, total
    words, n_commit[M], n_mem. The first n_state1 state entries are
    single-bit.
    Then, each section starting at an even word offset:
      inputs      n_in x {k, nbits, slot_a, slot_b}: bits of stimulus stream k
                  into slots slot_a ..; slot_b .. is the copy the clk = 1
                  settle reads
      state load  n_state x {cls, loc_lo, loc_hi, words, width, slot0}
      commit A0   n_commit[A0] x {state slot, next-state slot}: copies after
                  program A0 (next-state values that the program could not
                  write into their state slots directly)
      commit M    n_commit[M] x the same after program M
      phase A0, M, B: level offsets (nlev + 1 gate indices, padded to even),
                  then the gates as u64 (lo word first), evaluating
                  dst = s ? (x ^ ix) : (y ^ iy) on 64-lane planes; slot form
                  dst | s << 16 | x << 32 | ix << 47 | y << 48 | iy << 63,
                  or with byte_form (at most BYTE_FORM_MAX_SLOTS slots) the
                  LDS byte offsets 8 * slot, flags in bit 0:
                  8dst | 8s << 16 | (8x | ix) << 32 | (8y | iy) << 48
      hash M, hash B: n_hash x {nbytes, width, width slot refs} each
      memories    n_mem x {design memory index, cls, layout (1: lane-major),
                  loc_lo, loc_hi, row bytes, depth, width}
      ports A0, ports M: {words (including these two), n_ports}, then per
                  port, grouped by memory, reads before writes:
                  read  {0, mem, nabits, ntarget, nabits address refs,
                        ntarget x (data bit << 24 | state slot)}
                  write {1, mem, nabits, width, guard ref, nabits address
                        refs, width data refs}
      F chunks    n_fchunks x {nlev, ngates, n_wb, n_wb1, words, 0, level
                  offsets (padded to even), gates, write-back entries
                  {cls, loc_lo, loc_hi, words, width, width slot refs}, the
                  first n_wb1 single-bit (6 words each); ``words`` is the
                  chunk's size including its 6-word header}
    A slot ref is a slot index with the inversion in bit 31.
    """
    out: List[int] = [0] * HDR_WORDS
    byte_form = prog.nslots <= BYTE_FORM_MAX_SLOTS

    def pad(x):
        if len(x) & 1:
            x.append(0)

    def phase_words(ph) -> List[int]:
        w: List[int] = [0]
        for lev in ph.levels:
            w.append(w[-1] + len(lev))
        pad(w)
        for lev in ph.levels:
            for g in lev:
                gw = gate_word(g, byte_form)
                w += [gw & 0xffffffff, gw >> 32]
        return w

    pad(out)
    for k, nb, sa, sb in prog.in_chunks:
        out += [k, nb, sa, sb]
    pad(out)
    for cls, loc, words, width, s0 in prog.state_load:
        out += [cls, loc & 0xffffffff, loc >> 32, words, width, s0]
    for commit in prog.commits:
        pad(out)
        for st, dn in commit:
            out += [st, dn]
    for ph in prog.phases:
        pad(out)
        out += phase_words(ph)
    for hports in prog.hashes:
        pad(out)
        for nbytes, width, refs in hports:
            out += [nbytes, width] + refs
    pad(out)
    for idx, cls, layout, loc, row, depth, width in prog.mems:
        out += [idx, cls, layout, loc & 0xffffffff, loc >> 32, row, depth, width]
    for recs in prog.ports:
        pad(out)
        body: List[int] = []
        for r in recs:
            if r[0] == 0:
                _, mi, addr, targets = r
                body += [0, mi, len(addr), len(targets)] + addr + [(b << 24) | st for b, st in targets]
            else:
                _, mi, addr, g, data = r
                body += [1, mi, len(addr), len(data), g] + addr + data
        out += [2 + len(body), len(recs)] + body
    pad(out)
    for ph, entries in prog.fchunks:
        body = phase_words(ph) + _wb_words(entries)
        pad(body)
        n1 = sum(1 for e in entries if e[3] == 1)
        out += [len(ph.levels), sum(len(l) for l in ph.levels), len(entries), n1,
                FCHUNK_HDR + len(body), 0] + body
    pad(out)
    hdr = [BP_MAGIC, BP_VERSION | (int(byte_form) << 16), prog.nslots, len(prog.in_chunks),
           len(prog.state_load),
           len(prog.commits[0]), len(prog.hashes[0]), len(prog.fchunks)]
    hdr += [len(ph.levels) for ph in prog.phases]
    hdr += [sum(len(l) for l in ph.levels) for ph in prog.phases]
    hdr += [sum(1 for e in prog.state_load if e[3] == 1), len(out), len(prog.commits[1]),
            len(prog.mems)]
    out[:HDR_WORDS] = hdr
    return out


def sections(blob: Sequence[int]) -> Dict[str, object]:
    """Word offsets of every section of an encoded program (the runtime
    computes the same offsets from the header)."""
    h = blob[:HDR_WORDS]
    assert h[0] == BP_MAGIC and (h[1] & 0xffff) == BP_VERSION
    n_in, n_state, n_commit_a, n_hash, n_f = h[3:8]
    n_commit = (n_commit_a, h[16])
    nlev = h[8:11]
    ngates = h[11:14]
    o = HDR_WORDS
    ev = lambda x: x + (x & 1)
    res: Dict[str, object] = {}
    res["in"] = o = ev(o)
    o += 4 * n_in
    res["state"] = o = ev(o)
    o += 6 * n_state
    for k, name in enumerate(("commitA0", "commitM")):
        res[name] = o = ev(o)
        o += 2 * n_commit[k]
    for k, name in enumerate(("A0", "M", "B")):
        res["lev" + name] = o = ev(o)
        o += nlev[k] + 1
        res["gates" + name] = o = ev(o)
        o += 2 * ngates[k]
    for name in ("hashM", "hashB"):
        res[name] = o = ev(o)
        for _ in range(n_hash):
            o += 2 + blob[o + 1]
    res["mems"] = o = ev(o)
    o += 8 * h[17]
    for name in ("portsA0", "portsM"):
        res[name] = o = ev(o)
        o += blob[o]
    o = ev(o)
    chunks = []
    for _ in range(n_f):
        nl, ng, nwb, nwb1, words = blob[o:o + 5]
        lev = o + FCHUNK_HDR
        gates = ev(lev + nl + 1)
        chunks.append({"at": o, "nlev": nl, "ngates": ng, "n_wb": nwb, "n_wb1": nwb1, "lev": lev,
                       "gates": gates, "wb": gates + 2 * ng})
        o += words
    res["fchunks"] = chunks
    res["end"] = ev(o)
    assert res["end"] == h[15], (res["end"], h[15])
    return res
