"""Settle-loop and clocking coverage designs.

Combinational cycles through two always blocks (Verilator's UNOPTFLAT case):

* ``circ`` converges: ``y`` takes only the upper nibble of ``x`` and ``x``
  only the lower nibble of ``y`` (always 0), so after the clk = 0 settle
  ``x = a`` and ``y = a & 0xf0`` whatever the evaluation order, and the
  registered output is ``q = a & 0x0f`` of the previous cycle (closed form).
* ``osc`` never converges (``x = ~y``, ``y = x``): every eval must report
  non-convergence after the settle limit.
"""

CIRC = r"""
module circ(input clk, input [7:0] a, output [7:0] o, output [7:0] ox);
    reg [7:0] x, y, q;
    always @* x = a + (y & 8'h0f);
    always @* y = {x[7:4], 4'b0000};
    always @(posedge clk) q <= x ^ y;
    assign o = q;
    assign ox = x;
endmodule
"""

OSC = r"""
module osc(input clk, input a, output o);
    reg x, y, q;
    always @* x = ~y;
    always @* y = x ^ 1'b0;
    always @(posedge clk) q <= x ^ a;
    assign o = q;
endmodule
"""

# A gated clock: ``gclk = clk & en`` (en changes with the other inputs while
# clk is low), so ``g`` adds ``d`` exactly in the cycles whose ``en`` is set;
# ``c`` counts every cycle and ``s`` shifts in ``d[0]``.
GATED = r"""
module gated(input clk, input en, input [3:0] d, output [7:0] cnt, output [3:0] sr, output [7:0] gcnt);
    reg [7:0] c;
    reg [3:0] s;
    reg [7:0] g;
    wire gclk = clk & en;
    always @(posedge clk) begin
        c <= c + 8'd1;
        s <= {s[2:0], d[0]};
    end
    always @(posedge gclk) g <= g + {4'd0, d};
    assign cnt = c;
    assign sr = s;
    assign gcnt = g;
endmodule
"""


def gated_closed_form(stim, seed: int, sid: int, cycles: int) -> list:
    """Per-cycle output bytes (cnt, sr, gcnt) of ``gated`` for stimulus
    ``sid``; ``stim(seed, s, c, k)`` is the stimulus generator."""
    c = s = g = 0
    out = []
    for cy in range(cycles):
        en = stim(seed, sid, cy, 0) & 1
        d = stim(seed, sid, cy, 1 << 3) & 15
        c = (c + 1) & 255
        s = ((s << 1) | (d & 1)) & 15
        if en:
            g = (g + d) & 255
        out.append(bytes([c, s, g]))
    return out
