"""Known-answer designs built from the vectors of the reference's
self-checking arithmetic tests (tests/golden/kat_math.json, extracted by
tests/golden/make_kat.py from test_regress/t/t_math_shift.v, t_math_div.v,
t_math_signed.v and t_math_red.v).

Each design steps through its vectors, one per clock: a counter ``i``
selects the vector's operands and expected result from a case ROM, the
combinational logic computes every operation of the family, a selector
picks the vector's one, and the clocked block counts mismatches. Outputs:
``err`` (mismatches so far, 32 bits) and ``idx`` (vectors checked, 8 bits).
Every correct simulator therefore hashes the same sequence -- err = 0 and
idx = min(c + 1, N) after cycle c -- whatever the stimulus, which is the
closed form the tests compare every form against.
"""
from __future__ import annotations

import json
from pathlib import Path
from typing import List

KAT_JSON = Path(__file__).resolve().parents[2] / "tests" / "golden" / "kat_math.json"

SHIFT_OPS = ("left", "right", "rights")
DIV_OPS = ("qq", "rq", "qqs", "rqs", "qq4", "rq4")
RED_OPS = ("and_reduce", "or_reduce", "xor_reduce", "xnor_reduce")


def vectors() -> dict:
    return json.loads(KAT_JSON.read_text())


def _h(v: int, w: int) -> str:
    return f"{w}'h{v & ((1 << w) - 1):x}"


def _frame(name: str, n: int, decls: List[str], rom: List[str], logic: List[str]) -> str:
    L = [f"module {name}(input clk, input [7:0] nc, output [31:0] err, output [7:0] idx);",
         "    reg [7:0] i;",
         "    reg [31:0] e;"]
    L += ["    " + d for d in decls]
    L.append("    always @* begin")
    L.append("        case (i)")
    L += ["            " + r for r in rom]
    L.append("        endcase")
    L.append("    end")
    L += ["    " + x for x in logic]
    L.append("    always @(posedge clk) begin")
    L.append(f"        if (i < 8'd{n}) begin")
    L.append("            if (res != expv) e <= e + 32'd1;")
    L.append("            i <= i + 8'd1;")
    L.append("        end")
    L.append("    end")
    L.append("    assign err = e;")
    L.append("    assign idx = i;")
    L.append("endmodule")
    return "\n".join(L) + "\n"


def shift_verilog(vecs=None, name: str = "kat_shift") -> str:
    """Shifts left, right and arithmetic right of a 32-, 64- or 96-bit
    constant by a 32-, 64- or 96-bit amount (t_math_shift.v)."""
    vecs = vecs if vecs is not None else vectors()["shift"]
    consts = {}
    for v in vecs:
        consts[v["width"]] = v["value"]
    decls = ["reg [4:0] sel;", "reg [95:0] amt;", "reg [95:0] expv;", "reg [95:0] res;"]
    for w, c in sorted(consts.items()):
        decls.append(f"wire [{w - 1}:0] c{w} = {_h(c, w)};")
        decls.append(f"wire signed [{w - 1}:0] s{w} = {_h(c, w)};")
    decls += ["wire [31:0] a32 = amt[31:0];", "wire [63:0] a64 = amt[63:0];", "wire [95:0] a96 = amt;"]
    combos = [(op, w, aw) for op in SHIFT_OPS for w in sorted(consts) for aw in (32, 64, 96)]
    rom = []
    for k, v in enumerate(vecs):
        s = combos.index((v["op"], v["width"], v["amount_width"]))
        rom.append(f"8'd{k}: begin sel = 5'd{s}; amt = {_h(v['amount'], 96)}; "
                   f"expv = {_h(v['expect'], 96)}; end")
    rom.append("default: begin sel = 5'd0; amt = 96'd0; expv = 96'd0; end")
    logic = ["always @* begin", "    case (sel)"]
    for s, (op, w, aw) in enumerate(combos):
        x = {"left": f"c{w} << a{aw}", "right": f"c{w} >> a{aw}", "rights": f"s{w} >>> a{aw}"}[op]
        pad = 96 - w
        val = f"{{{pad}'d0, {x}}}" if pad else f"({x})"
        logic.append(f"        5'd{s}: res = {val};")
    logic += ["        default: res = 96'd0;", "    endcase", "end"]
    return _frame(name, len(vecs), decls, rom, logic)


def div_verilog(vecs=None, name: str = "kat_div") -> str:
    """61-bit unsigned and signed quotient and remainder, and by the
    constant 4 (t_math_div.v); division by zero gives 0."""
    vecs = vecs if vecs is not None else vectors()["div"]
    decls = ["reg [2:0] sel;", "reg [60:0] a;", "reg [60:0] b;", "reg [60:0] expv;", "reg [60:0] res;",
             "wire [60:0] qq = a / b;", "wire [60:0] rq = a % b;",
             "wire [60:0] qqs = $signed(a) / $signed(b);", "wire [60:0] rqs = $signed(a) % $signed(b);",
             "wire [60:0] qq4 = a / 4;", "wire [60:0] rq4 = a % 4;"]
    rom = []
    for k, v in enumerate(vecs):
        rom.append(f"8'd{k}: begin sel = 3'd{DIV_OPS.index(v['op'])}; a = {_h(v['a'], 61)}; "
                   f"b = {_h(v['b'], 61)}; expv = {_h(v['expect'], 61)}; end")
    rom.append("default: begin sel = 3'd0; a = 61'd0; b = 61'd1; expv = 61'd0; end")
    logic = ["always @* begin", "    case (sel)"]
    for s, op in enumerate(DIV_OPS):
        logic.append(f"        3'd{s}: res = {op};")
    logic += ["        default: res = 61'd0;", "    endcase", "end"]
    return _frame(name, len(vecs), decls, rom, logic)


def signed_verilog(vecs=None, name: str = "kat_signed") -> str:
    """Shifts of a signed 16-bit value by a signed 5-bit amount and the sign
    extension rules of t_math_signed.v: {a >> b, a >>> b, a << b, a <<< b,
    b >>> 4 (b signed, extended to 16 bits), b[4:0] >>> 4 (a part-select
    is unsigned)}."""
    vecs = vecs if vecs is not None else vectors()["signed"]
    decls = ["reg signed [15:0] a;", "reg signed [4:0] b;", "reg [95:0] expv;",
             "reg signed [15:0] sr, srs, sl, sls;", "reg [15:0] b_s, b_us;", "reg [95:0] res;"]
    rom = []
    for k, v in enumerate(vecs):
        rom.append(f"8'd{k}: begin a = {_h(v['a'], 16)}; b = {_h(v['b'], 5)}; "
                   f"expv = {_h(v['expect'], 96)}; end")
    rom.append("default: begin a = 16'd0; b = 5'd0; expv = 96'd0; end")
    logic = ["always @* begin",
             "    sr = a >> b;",
             "    srs = a >>> b;",
             "    sl = a << b;",
             "    sls = a <<< b;",
             "    b_s = b >>> 4;",
             "    b_us = b[4:0] >>> 4;",
             "    res = {sr, srs, sl, sls, b_s, b_us};",
             "end"]
    return _frame(name, len(vecs), decls, rom, logic)


def red_verilog(vecs=None, name: str = "kat_red") -> str:
    """AND, OR, XOR and XNOR reductions of a 68-bit value (t_math_red.v)."""
    vecs = vecs if vecs is not None else vectors()["red"]
    decls = ["reg [1:0] sel;", "reg [67:0] r;", "reg expv;", "reg res;"]
    rom = []
    for k, v in enumerate(vecs):
        rom.append(f"8'd{k}: begin sel = 2'd{RED_OPS.index(v['op'])}; r = {_h(v['value'], 68)}; "
                   f"expv = 1'b{v['expect']}; end")
    rom.append("default: begin sel = 2'd0; r = 68'd0; expv = 1'b0; end")
    logic = ["always @* begin", "    case (sel)",
             "        2'd0: res = &r;", "        2'd1: res = |r;", "        2'd2: res = ^r;",
             "        default: res = ~^r;", "    endcase", "end"]
    return _frame(name, len(vecs), decls, rom, logic)


FAMILIES = {"shift": shift_verilog, "div": div_verilog, "signed": signed_verilog, "red": red_verilog}


def verilog(family: str, vecs=None) -> str:
    return FAMILIES[family](vecs, name=f"kat_{family}")


def expected_hashes(fnv, n_vectors: int, cycles: int, nstim: int) -> list:
    """The closed form: err = 0 and idx = min(c + 1, N) after each cycle,
    the same for every stimulus. ``fnv(data, h)`` is FNV-1a-64."""
    h = 0xCBF29CE484222325
    for c in range(cycles):
        h = fnv((0).to_bytes(4, "little") + bytes([min(c + 1, n_vectors)]), h)
    return [h] * nstim
