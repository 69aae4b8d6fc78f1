"""Synthetic design 4: a 16 x 16 systolic array of int8 x int8 MAC processing
elements with 32-bit accumulators, plus a 4,096-bit shift/compare chain.

Inputs per cycle: ``row_in`` (128 bits: one int8 per row, row i in bits
[8i+7:8i]) and ``col_in`` (128 bits: one int8 per column).

Processing element (i, j) holds a (int8), b (int8) and acc (int32):
  a(i, j) <= j == 0 ? row_in[i] : a(i, j-1)      (operands flow right)
  b(i, j) <= i == 0 ? col_in[j] : b(i-1, j)      (operands flow down)
  acc(i, j) <= acc(i, j) + a(i, j) * b(i, j)     (signed, modulo 2^32)
all from the values before the edge.

Chain: ``chain`` (4,096 bits) <= {chain[4087:0], row_in[7:0] ^ col_in[7:0]};
``cmp`` <= chain[4095:2048] < chain[2047:0] (unsigned); ``par`` <= ^chain.

Outputs: ``acc_x`` = XOR of all 256 accumulators, ``flags`` = {cmp, par}.
"""
from __future__ import annotations

N = 16


def verilog(n: int = N, chain_bits: int = 4096) -> str:
    L = []
    w = L.append
    w("module systolic(")
    w("    input clk,")
    w(f"    input [{8 * n - 1}:0] row_in,")
    w(f"    input [{8 * n - 1}:0] col_in,")
    w("    output [31:0] acc_x,")
    w("    output [1:0] flags")
    w(");")
    for i in range(n):
        for j in range(n):
            w(f"    reg signed [7:0] a_{i}_{j};")
            w(f"    reg signed [7:0] b_{i}_{j};")
            w(f"    reg signed [31:0] acc_{i}_{j};")
    w(f"    reg [{chain_bits - 1}:0] chain;")
    w("    reg cmp, par;")
    w("    always @(posedge clk) begin")
    for i in range(n):
        for j in range(n):
            a_src = f"row_in[{8 * i + 7}:{8 * i}]" if j == 0 else f"a_{i}_{j - 1}"
            b_src = f"col_in[{8 * j + 7}:{8 * j}]" if i == 0 else f"b_{i - 1}_{j}"
            w(f"        a_{i}_{j} <= {a_src};")
            w(f"        b_{i}_{j} <= {b_src};")
            w(f"        acc_{i}_{j} <= acc_{i}_{j} + a_{i}_{j} * b_{i}_{j};")
    w(f"        chain <= {{chain[{chain_bits - 9}:0], row_in[7:0] ^ col_in[7:0]}};")
    half = chain_bits // 2
    w(f"        cmp <= chain[{chain_bits - 1}:{half}] < chain[{half - 1}:0];")
    w("        par <= ^chain;")
    w("    end")
    terms = " ^ ".join(f"acc_{i}_{j}" for i in range(n) for j in range(n))
    w(f"    assign acc_x = {terms};")
    w("    assign flags = {cmp, par};")
    w("endmodule")
    return "\n".join(L) + "\n"
