"""Reference interpreter of the design IR (TEST INFRASTRUCTURE ONLY).

Evaluates an elaborated design for a handful of stimuli with Python integers
and the clock-cycle protocol, independently of the compiler's ordering,
partitioning, storage and emitted code. Only small cases (pure-Python loops).

Semantics restated (two-state Verilog as Verilator computes it, with
``--x-assign 0 --x-initial 0``):

* values are unsigned bit patterns of their width; signed operators interpret
  the operand's top bit as the sign (include/rf_verilated.h:1761-1817);
* shifts by at least the width give 0 (or all sign bits for >>>), selects and
  memory reads out of range give 0, writes out of range are dropped, x / 0 = 0;
* an eval: the combinational logic settles on the new inputs (so a clock
  derived from them, such as a gated clock, takes its new value), clocked
  blocks whose edge fired run against the values from before the edge
  (non-blocking assignments committed after all of them), then the
  combinational logic is iterated to a fixpoint again (the settle loop;
  acyclic logic reaches it after one ordered pass), then every clock's
  latch, derived clocks included, is updated.

Op semantics here are written from the definitions above, not shared with the
compiler's constant folder.
"""
from __future__ import annotations

import sys
from pathlib import Path
from typing import Dict, List

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from rtlflow_amd import ir  # noqa: E402  (IR node classes only)

M64 = (1 << 64) - 1


def _mask(w):
    return (1 << w) - 1


def _sgn(v, w):
    return v - (1 << w) if (v >> (w - 1)) & 1 else v


def _splitmix(x):
    z = (x + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _stim(seed, s, c, k):
    return _splitmix(seed ^ ((s * 0x9E3779B97F4A7C15) & M64) ^ ((c << 8) & M64) ^ k)


def _fnv(h, data: bytes):
    for b in data:
        h ^= b
        h = (h * 0x100000001b3) & M64
    return h


class Lane:
    def __init__(self, design: ir.Design):
        self.d = design
        self.v: Dict[ir.Var, int] = {}
        self.m: Dict[ir.Var, List[int]] = {}
        for var in design.vars:
            if var.is_mem:
                self.m[var] = [0] * var.depth
            else:
                self.v[var] = 0
        self.clklast = {c: 0 for c in design.clocks}
        self.did_init = False

    # ---------------------------------------------------------- expressions
    def ev(self, e) -> int:
        w = e.width
        if isinstance(e, ir.Const):
            return e.value
        if isinstance(e, ir.Ref):
            return self.v[e.var]
        if isinstance(e, ir.MemRd):
            i = self.ev(e.idx)
            mem = self.m[e.var]
            return mem[i] if i < len(mem) else 0
        if isinstance(e, ir.Sel):
            a = self.ev(e.a)
            lsb = e.lsb if isinstance(e.lsb, int) else self.ev(e.lsb)
            if lsb >= e.a.width:
                return 0
            return (a >> lsb) & _mask(w)
        if isinstance(e, ir.Concat):
            r = 0
            for p in e.parts:
                r = (r << p.width) | self.ev(p)
            return r
        if isinstance(e, ir.Cond):
            return self.ev(e.a) if self.ev(e.c) else self.ev(e.b)
        if isinstance(e, ir.Op):
            return self.op(e)
        raise TypeError(e)

    def op(self, e: ir.Op) -> int:
        o, w = e.op, e.width
        xs = [self.ev(a) for a in e.args]
        wa = e.args[0].width
        a = xs[0]
        if o == "not":
            return a ^ _mask(w)
        if o == "neg":
            return (-a) & _mask(w)
        if o == "zext":
            return a
        if o == "sext":
            return _sgn(a, wa) & _mask(w)
        if o == "redand":
            return int(a == _mask(wa))
        if o == "redor":
            return int(a != 0)
        if o == "redxor":
            return bin(a).count("1") % 2
        if o == "lognot":
            return int(a == 0)
        b = xs[1]
        if o in ("shl", "shr", "sra"):
            if o == "shl":
                return (a << b) & _mask(w) if b < w else 0
            if o == "shr":
                return a >> b if b < w else 0
            return (_sgn(a, w) >> min(b, w)) & _mask(w)
        if o == "logand":
            return int(bool(a) and bool(b))
        if o == "logor":
            return int(bool(a) or bool(b))
        if o in ("eq", "ne", "ltu", "leu", "gtu", "geu"):
            return int({"eq": a == b, "ne": a != b, "ltu": a < b, "leu": a <= b, "gtu": a > b,
                        "geu": a >= b}[o])
        if o in ("lts", "les", "gts", "ges"):
            sa, sb = _sgn(a, wa), _sgn(b, wa)
            return int({"lts": sa < sb, "les": sa <= sb, "gts": sa > sb, "ges": sa >= sb}[o])
        m = _mask(w)
        if o == "and":
            return a & b
        if o == "or":
            return a | b
        if o == "xor":
            return a ^ b
        if o == "xnor":
            return (a ^ b) ^ m
        if o == "add":
            return (a + b) & m
        if o == "sub":
            return (a - b) & m
        if o == "mul":
            return (a * b) & m
        if o == "divu":
            return a // b if b else 0
        if o == "modu":
            return a % b if b else 0
        if o in ("divs", "mods"):
            if b == 0:
                return 0
            sa, sb = _sgn(a, w), _sgn(b, w)
            q = abs(sa) // abs(sb) * (1 if (sa < 0) == (sb < 0) else -1)
            return (q if o == "divs" else sa - q * sb) & m
        raise ValueError(o)

    # ----------------------------------------------------------- statements
    def run(self, stmts, nba: list, is_seq: bool):
        for s in stmts:
            if isinstance(s, ir.Assign):
                val = self.ev(s.rhs)
                tgt = s.lhs
                idx = self.ev(tgt.idx) if tgt.idx is not None else None
                lsb = None
                if tgt.lsb is not None:
                    lsb = tgt.lsb if isinstance(tgt.lsb, int) else self.ev(tgt.lsb)
                if s.nb:
                    nba.append((tgt, idx, lsb, val))
                else:
                    self.write(tgt, idx, lsb, val)
            elif isinstance(s, ir.If):
                self.run(s.then if self.ev(s.cond) else s.els, nba, is_seq)
            elif isinstance(s, ir.Case):
                sel = self.ev(s.expr)
                done = False
                for labels, body in s.items:
                    if labels is None:
                        continue
                    if any(((sel ^ v) & care) == 0 for v, care in labels):
                        self.run(body, nba, is_seq)
                        done = True
                        break
                if not done:
                    for labels, body in s.items:
                        if labels is None:
                            self.run(body, nba, is_seq)
                            break

    def write(self, tgt: ir.Lhs, idx, lsb, val):
        var = tgt.var
        if var.is_mem:
            mem = self.m[var]
            if idx >= len(mem):
                return
            if lsb is None:
                mem[idx] = val
            else:
                mem[idx] = self.insert(mem[idx], var.width, lsb, tgt.width, val)
            return
        if lsb is None:
            self.v[var] = val
        else:
            self.v[var] = self.insert(self.v[var], var.width, lsb, tgt.width, val)

    @staticmethod
    def insert(old, wv, lsb, wl, val):
        if lsb >= wv:
            return old
        m = (_mask(wl) << lsb) & _mask(wv)
        return (old & ~m) | ((val << lsb) & m)

    # ------------------------------------------------------------------ eval
    def comb_fixpoint(self, limit=100):
        combs = [p for p in self.d.processes if p.kind == "comb"]
        for _ in range(limit * max(1, len(combs))):
            before = (dict(self.v), {k: list(x) for k, x in self.m.items()})
            for p in combs:
                self.run(p.stmts, [], False)
            if before == (self.v, self.m):
                return
        raise RuntimeError("combinational logic did not converge")

    def eval(self):
        if not self.did_init:
            self.did_init = True
            for p in self.d.processes:
                if p.kind == "initial":
                    nba = []
                    self.run(p.stmts, nba, True)
                    for t in nba:
                        self.write(*t)
            self.comb_fixpoint()
            for c in self.d.clocks:
                self.clklast[c] = self.v[c]
        else:
            self.comb_fixpoint()
        fired = []
        for p in self.d.processes:
            if p.kind != "seq":
                continue
            for kind, clk in p.edges:
                now, last = self.v[clk], self.clklast[clk]
                if (kind == "posedge" and now and not last) or (kind == "negedge" and not now and last):
                    fired.append(p)
                    break
        nba = []
        for p in fired:
            self.run(p.stmts, nba, True)
        for t in nba:
            self.write(*t)
        self.comb_fixpoint()
        for c in self.d.clocks:
            self.clklast[c] = self.v[c]


def class_bytes(w):
    return 1 if w <= 8 else 2 if w <= 16 else 4 if w <= 32 else 8 if w <= 64 else 4 * ((w + 31) // 32)


def protocol(design: ir.Design, seed: int, nstim: int, cycles: int, stim0: int = 0,
             images=None, budgets=None, inputs=None, capture=None) -> List[int]:
    """Per-stimulus FNV-1a-64 of the outputs under the clock-cycle protocol.
    ``images``: {memory name: per-stimulus initial contents, indexed by
    stimulus id}, loaded before the initial blocks run; ``budgets``: cycles
    per stimulus (indexed from stim0) instead of ``cycles``. ``inputs``: a
    [cycle][word][stimulus - stim0] array of 64-bit input words (data inputs
    in declaration order, 64 bits per word) used instead of the generator;
    ``capture``: an array of the same layout over the output words, filled
    with every cycle's outputs."""
    data_inputs = [v for v in design.vars if v.kind == "input" and v not in design.clocks]
    outputs = [v for v in design.vars if v.kind == "output"]
    mems = {v.name: v for v in design.vars if v.is_mem}
    hashes = []
    for s in range(stim0, stim0 + nstim):
        lane = Lane(design)
        for name, img in (images or {}).items():
            row = [int(x) for x in img[s]]
            mem = lane.m[mems[name]]
            mem[:len(row)] = row[:len(mem)]
        h = 0xcbf29ce484222325
        ncyc = cycles if budgets is None else int(budgets[s - stim0])
        for c in range(ncyc):
            n = 0
            for p, var in enumerate(data_inputs):
                val = 0
                for j in range((var.width + 63) // 64):
                    w = _stim(seed, s, c, (p << 3) | j) if inputs is None else int(inputs[c][n][s - stim0])
                    val |= w << (64 * j)
                    n += 1
                lane.v[var] = val & _mask(var.width)
            for ph in (0, 1):
                for clk in design.clocks:
                    lane.v[clk] = ph
                lane.eval()
            n = 0
            for var in outputs:
                h = _fnv(h, lane.v[var].to_bytes(class_bytes(var.width), "little"))
                if capture is not None:
                    for j in range((var.width + 63) // 64):
                        capture[c][n][s - stim0] = (lane.v[var] >> (64 * j)) & 0xFFFFFFFFFFFFFFFF
                        n += 1
        hashes.append(h)
    return hashes
