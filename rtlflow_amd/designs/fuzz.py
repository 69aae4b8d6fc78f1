"""Seeded random Verilog designs for op-semantics coverage.

Each design has inputs of assorted widths (1 to 130 bits, some signed), random
combinational expressions over them (every operator the front end supports,
constant and dynamic selects, concatenation and replication, signed casts,
shifts by narrow and wide amounts, wide arithmetic and compares), clocked
registers with conditional and partial non-blocking writes, a small memory,
and one output that concatenates a sample of everything. They are compared
against the IR interpreter in oracle/ (CPU) and on the GPU.
"""
from __future__ import annotations

import random
from typing import List, Tuple

WIDTHS = [1, 3, 7, 8, 13, 16, 31, 32, 33, 47, 63, 64, 65, 96, 128, 130]


class Gen:
    def __init__(self, seed: int, wide: bool = True, bitlevel: bool = False):
        self.bitlevel = bitlevel
        self.r = random.Random(seed)
        self.wide = wide
        self.sigs: List[Tuple[str, int, bool]] = []   # name, width, signed

    def width(self) -> int:
        ws = WIDTHS if self.wide else [w for w in WIDTHS if w <= 64]
        return self.r.choice(ws)

    def leaf(self) -> Tuple[str, int]:
        r = self.r
        k = r.random()
        if k < 0.15:
            w = self.width()
            v = r.getrandbits(w)
            return f"{w}'h{v:x}", w
        if k < 0.2:
            return str(r.randint(0, 1000)), 32
        name, w, s = r.choice(self.sigs)
        k = r.random()
        if k < 0.25 and w > 1:
            hi = r.randint(0, w - 1)
            lo = r.randint(0, hi)
            return f"{name}[{hi}:{lo}]", hi - lo + 1
        if k < 0.33 and w > 1:
            idx, _ = self.narrow_expr(2)
            return f"{name}[{idx}]", 1
        if k < 0.4 and w > 4:
            pw = r.randint(1, min(w, 16))
            st, _ = self.narrow_expr(2)
            if self.bitlevel:
                return f"{name}[{st} +: {pw}]", pw
            return f"{name}[({st}) % {w} +: {pw}]", pw
        if s:
            return f"$signed({name})", w
        return name, w

    def narrow_expr(self, depth: int) -> Tuple[str, int]:
        name, w, _ = self.r.choice([x for x in self.sigs if x[1] <= 16] or self.sigs)
        if self.r.random() < 0.5 and w > 3:
            return f"{name}[2:0]", 3
        return name, w

    def expr(self, depth: int) -> Tuple[str, int]:
        r = self.r
        if depth <= 0 or r.random() < 0.2:
            return self.leaf()
        k = r.random()
        if k < 0.35:
            op = r.choice(["+", "-", "*", "&", "|", "^", "~^"])
            a, wa = self.expr(depth - 1)
            b, wb = self.expr(depth - 1)
            if op == "*" and max(wa, wb) > 64 and r.random() < 0.5:
                op = "+"
            if op == "*" and self.bitlevel:
                op = "&"
            return f"({a} {op} {b})", max(wa, wb)
        if k < 0.42:
            op = r.choice(["/", "%"])
            a, wa = self.expr(depth - 1)
            b, wb = self.expr(depth - 1)
            if max(wa, wb) > 64 or self.bitlevel:
                return f"({a} ^ {b})", max(wa, wb)
            return f"({a} {op} {b})", max(wa, wb)
        if k < 0.52:
            op = r.choice(["==", "!=", "<", "<=", ">", ">="])
            a, wa = self.expr(depth - 1)
            b, wb = self.expr(depth - 1)
            return f"({a} {op} {b})", 1
        if k < 0.6:
            op = r.choice(["<<", ">>", ">>>", "<<<"])
            a, wa = self.expr(depth - 1)
            if r.random() < 0.2:
                b, _ = self.expr(depth - 1)
            else:
                b, _ = self.narrow_expr(1)
            return f"({a} {op} {b})", wa
        if k < 0.68:
            op = r.choice(["~", "-", "!", "&", "|", "^", "~&", "~|", "~^"])
            a, wa = self.expr(depth - 1)
            return f"({op}{a})", (wa if op in ("~", "-") else 1)
        if k < 0.76:
            c, _ = self.expr(depth - 1)
            a, wa = self.expr(depth - 1)
            b, wb = self.expr(depth - 1)
            return f"({c} ? {a} : {b})", max(wa, wb)
        if k < 0.84:
            n = r.randint(2, 3)
            parts = [self.expr(depth - 1) for _ in range(n)]
            return "{" + ", ".join(p for p, _ in parts) + "}", sum(w for _, w in parts)
        if k < 0.88:
            a, wa = self.leaf()
            n = r.randint(2, 4)
            return f"{{{n}{{{a}}}}}", n * wa
        if k < 0.94:
            a, wa = self.expr(depth - 1)
            return f"$signed({a})", wa
        if k < 0.97:
            a, wa = self.expr(depth - 1)
            b, wb = self.expr(depth - 1)
            return f"({a} {r.choice(['&&', '||'])} {b})", 1
        a, wa = self.expr(depth - 1)
        return f"$unsigned({a})", wa


def verilog(seed: int, n_in: int = 5, n_wire: int = 8, n_reg: int = 6, depth: int = 3,
            wide: bool = True, name: str = "fuzz", bitlevel: bool = False) -> str:
    """``bitlevel``: no memory, no division, multiplies at most 16 bits wide
    (the subset the bit-plane form lowers)."""
    g = Gen(seed, wide, bitlevel)
    r = g.r
    L = []
    ports = ["input clk"]
    for i in range(n_in):
        w = g.width()
        s = r.random() < 0.3
        ports.append(f"input {'signed ' if s else ''}[{w - 1}:0] i{i}")
        g.sigs.append((f"i{i}", w, s))
    body = []
    regs = []
    for i in range(n_reg):
        w = g.width()
        s = r.random() < 0.3
        body.append(f"    reg {'signed ' if s else ''}[{w - 1}:0] r{i};")
        regs.append((f"r{i}", w, s))
    g.sigs.extend(regs)
    mw = g.width() if wide else r.choice([8, 16, 32, 64])
    if not bitlevel:
        body.append(f"    reg [{mw - 1}:0] mem [0:15];")
    for i in range(n_wire):
        e, w = g.expr(depth)
        wdecl = r.choice([w, g.width()])
        s = r.random() < 0.2
        body.append(f"    wire {'signed ' if s else ''}[{wdecl - 1}:0] w{i} = {e};")
        g.sigs.append((f"w{i}", wdecl, s))
    if not bitlevel:
        body.append(f"    wire [{mw - 1}:0] mrd = mem[w0[3:0]];")
        g.sigs.append(("mrd", mw, False))
    body.append("    always @(posedge clk) begin")
    for name_, w, _ in regs:
        e, _ = g.expr(depth)
        k = r.random()
        if k < 0.3:
            c, _ = g.expr(1)
            e2, _ = g.expr(depth - 1)
            body.append(f"        if ({c}) {name_} <= {e}; else {name_} <= {e2};")
        elif k < 0.5 and w > 2:
            hi = r.randint(1, w - 1)
            lo = r.randint(0, hi)
            body.append(f"        {name_}[{hi}:{lo}] <= {e};")
        elif k < 0.6 and w > 2:
            c, _ = g.narrow_expr(1)
            body.append(f"        case ({c})")
            body.append(f"            3'd1, 3'd5: {name_} <= {e};")
            body.append(f"            3'd2: {name_} <= {name_} + 1;")
            body.append("            default: ;")
            body.append("        endcase")
        else:
            body.append(f"        {name_} <= {e};")
    e, _ = g.expr(2)
    if not bitlevel:
        body.append(f"        if (w1[0]) mem[i0[3:0]] <= {e};")
    body.append("    end")
    picks = [s for s in g.sigs if not s[0].startswith("i")]
    outs = r.sample(picks, min(6, len(picks)))
    ow = sum(w for _, w, _ in outs)
    ports.append(f"output [{ow - 1}:0] out")
    body.append("    assign out = {" + ", ".join(n for n, _, _ in outs) + "};")
    L.append(f"module {name}(" + ", ".join(ports) + ");")
    L.extend(body)
    L.append("endmodule")
    return "\n".join(L) + "\n"
