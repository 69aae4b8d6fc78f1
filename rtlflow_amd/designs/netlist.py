"""Seeded random gate-level netlist generator (synthetic designs 2 and 5).

Structure: ``n_gates`` gates in an even mix of AND / OR / XOR / NAND / 2:1 MUX,
``n_ff`` D flip-flops on one clock, ``n_pi`` primary inputs (one input port
``pi``), ``n_po`` primary outputs (one output port ``po``). Gates are spread
evenly over ``depth`` levels; each fan-in is drawn uniformly from all nodes of
earlier levels (level 0 = primary inputs and flip-flop outputs), so the
combinational depth is at most ``depth``. Flip-flop D inputs and primary
outputs are drawn uniformly from the gates. Optional synchronous RAMs
(``n_rams`` x ``ram_depth`` words of ``ram_width`` bits): each has an address,
write-enable and write-data drawn from gates, a registered read port feeding
level 0, and the read data's bits join the fan-in pool.

Two artefacts with the same content: structural Verilog for the compiler and a
compact gate table (``GateTable``) for the independent C evaluator in
``oracle/``; the oracle never sees the Verilog.
"""
from __future__ import annotations

import io
import struct
from dataclasses import dataclass, field
from typing import List

import numpy as np

AND, OR, XOR, NAND, MUX = 0, 1, 2, 3, 4
OPNAME = {AND: "&", OR: "|", XOR: "^"}

_M64 = (1 << 64) - 1


class SplitMix:
    """Generator stream (SplitMix64) for design construction."""

    def __init__(self, seed: int):
        self.s = seed & _M64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & _M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def below(self, n: int) -> int:
        return self.next() % n


@dataclass
class GateTable:
    """Node ids: [0, n_pi) primary-input bits, [n_pi, n_pi+n_ff) flip-flop
    outputs, then RAM read-data bits, then gates in level order."""
    n_pi: int
    n_ff: int
    n_ram_bits: int
    gate_type: np.ndarray   # uint8 [n_gates]
    gate_in: np.ndarray     # uint32 [n_gates, 3] (third used by MUX: sel ? a : b)
    ff_d: np.ndarray        # uint32 [n_ff]
    po: np.ndarray          # uint32 [n_po]
    # RAMs: per RAM: address node list (lsb first), we node, data nodes (lsb first)
    ram_depth: int = 0
    ram_width: int = 0
    ram_addr: np.ndarray = field(default_factory=lambda: np.zeros((0, 0), np.uint32))
    ram_we: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.uint32))
    ram_wdata: np.ndarray = field(default_factory=lambda: np.zeros((0, 0), np.uint32))

    @property
    def n_gates(self) -> int:
        return len(self.gate_type)

    @property
    def first_gate(self) -> int:
        return self.n_pi + self.n_ff + self.n_ram_bits

    def to_bytes(self) -> bytes:
        """Binary format read by oracle/netlist_eval.c (little-endian u32s):
        magic 'RFNL', n_pi, n_ff, n_ram_bits, n_gates, n_po, n_rams, ram_depth,
        ram_width, addr_bits, then gate_type[n_gates] (u32 each),
        gate_in[n_gates*3], ff_d[n_ff], po[n_po], and per RAM: addr[addr_bits],
        we, wdata[ram_width]."""
        n_rams = len(self.ram_we)
        abits = self.ram_addr.shape[1] if n_rams else 0
        hdr = struct.pack("<4s9I", b"RFNL", self.n_pi, self.n_ff, self.n_ram_bits, self.n_gates,
                          len(self.po), n_rams, self.ram_depth, self.ram_width, abits)
        out = io.BytesIO()
        out.write(hdr)
        out.write(self.gate_type.astype("<u4").tobytes())
        out.write(self.gate_in.astype("<u4").tobytes())
        out.write(self.ff_d.astype("<u4").tobytes())
        out.write(self.po.astype("<u4").tobytes())
        for r in range(n_rams):
            out.write(self.ram_addr[r].astype("<u4").tobytes())
            out.write(struct.pack("<I", int(self.ram_we[r])))
            out.write(self.ram_wdata[r].astype("<u4").tobytes())
        return out.getvalue()


def generate(n_gates: int = 20000, n_ff: int = 2000, n_pi: int = 64, n_po: int = 64,
             depth: int = 40, seed: int = 0x52544C4600000002, n_rams: int = 0,
             ram_depth: int = 1024, ram_width: int = 64, name: str = "netlist"):
    """Returns (verilog_text, GateTable)."""
    rng = SplitMix(seed)
    n_ram_bits = n_rams * ram_width
    first_gate = n_pi + n_ff + n_ram_bits
    level0 = first_gate
    per_level = [n_gates // depth + (1 if i < n_gates % depth else 0) for i in range(depth)]
    gate_type = np.zeros(n_gates, np.uint8)
    gate_in = np.zeros((n_gates, 3), np.uint32)
    g = 0
    pool = level0          # nodes [0, pool) are available to the current level
    for lvl in range(depth):
        for _ in range(per_level[lvl]):
            t = g % 5
            gate_type[g] = t
            for k in range(3 if t == MUX else 2):
                gate_in[g, k] = rng.below(pool)
            g += 1
        pool = first_gate + g
    total = first_gate + n_gates
    ff_d = np.array([first_gate + rng.below(n_gates) for _ in range(n_ff)], np.uint32)
    po = np.array([first_gate + rng.below(n_gates) for _ in range(n_po)], np.uint32)
    abits = max(1, (ram_depth - 1).bit_length()) if n_rams else 0
    ram_addr = np.array([[first_gate + rng.below(n_gates) for _ in range(abits)]
                         for _ in range(n_rams)], np.uint32).reshape(n_rams, abits)
    ram_we = np.array([first_gate + rng.below(n_gates) for _ in range(n_rams)], np.uint32)
    ram_wdata = np.array([[first_gate + rng.below(n_gates) for _ in range(ram_width)]
                          for _ in range(n_rams)], np.uint32).reshape(n_rams, ram_width)
    table = GateTable(n_pi, n_ff, n_ram_bits, gate_type, gate_in, ff_d, po, ram_depth, ram_width,
                      ram_addr, ram_we, ram_wdata)
    return to_verilog(table, name), table


def node_name(t: GateTable, i: int) -> str:
    if i < t.n_pi:
        return f"pi[{i}]"
    i -= t.n_pi
    if i < t.n_ff:
        return f"q{i}"
    i -= t.n_ff
    if i < t.n_ram_bits:
        return f"rd{i // t.ram_width}[{i % t.ram_width}]"
    return f"n{i - t.n_ram_bits}"


def to_verilog(t: GateTable, name: str = "netlist") -> str:
    out = io.StringIO()
    w = out.write
    w(f"// seeded random netlist: {t.n_gates} gates, {t.n_ff} flip-flops\n")
    w(f"module {name}(input clk, input [{t.n_pi - 1}:0] pi, output [{len(t.po) - 1}:0] po);\n")
    for i in range(t.n_ff):
        w(f"  reg q{i};\n")
    n_rams = len(t.ram_we)
    abits = t.ram_addr.shape[1] if n_rams else 0
    for r in range(n_rams):
        w(f"  reg [{t.ram_width - 1}:0] ram{r} [0:{t.ram_depth - 1}];\n")
        w(f"  reg [{t.ram_width - 1}:0] rd{r};\n")
    for g in range(t.n_gates):
        w(f"  wire n{g};\n")
    nn = lambda i: node_name(t, int(i))
    for g in range(t.n_gates):
        ty = int(t.gate_type[g])
        a, b, c = (nn(x) for x in t.gate_in[g])
        if ty == MUX:
            w(f"  assign n{g} = {a} ? {b} : {c};\n")
        elif ty == NAND:
            w(f"  assign n{g} = ~({a} & {b});\n")
        else:
            w(f"  assign n{g} = {a} {OPNAME[ty]} {b};\n")
    w("  always @(posedge clk) begin\n")
    for i in range(t.n_ff):
        w(f"    q{i} <= {nn(t.ff_d[i])};\n")
    w("  end\n")
    for r in range(n_rams):
        addr = "{" + ", ".join(nn(x) for x in reversed(t.ram_addr[r])) + "}"
        data = "{" + ", ".join(nn(x) for x in reversed(t.ram_wdata[r])) + "}"
        w("  always @(posedge clk) begin\n")
        w(f"    if ({nn(t.ram_we[r])}) ram{r}[{addr}] <= {data};\n")
        w(f"    rd{r} <= ram{r}[{addr}];\n")
        w("  end\n")
    w("  assign po = {" + ", ".join(nn(x) for x in reversed(t.po)) + "};\n")
    w("endmodule\n")
    return out.getvalue()
