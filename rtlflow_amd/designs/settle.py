"""Settle-loop and clocking coverage designs.

Combinational cycles through two always blocks (Verilator's UNOPTFLAT case):

* ``circ`` converges: ``y`` takes only the upper nibble of ``x`` and ``x``
  only the lower nibble of ``y`` (always 0), so after the clk = 0 settle
  ``x = a`` and ``y = a & 0xf0`` whatever the evaluation order, and the
  registered output is ``q = a & 0x0f`` of the previous cycle (closed form).
* ``smear`` converges after up to 17 passes (see below).
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

# ``smear`` needs many settle passes: ``x = (1 << a) | (y << 1)`` with
# ``y = x`` through two always blocks is a combinational cycle whose unique
# fixed point is ``x = 0xffff << a`` (bit i of x is set iff some bit j <= i
# of ``1 << a`` is). Each pass moves a change one bit further, so settling
# from the previous cycle's fixed point (shift p) to this cycle's (shift q)
# takes about q - p + 1 passes and the first cycle from reset up to 17: more
# than the graph form's captured passes, fewer than the converge limit. The
# register ``r`` sums the settled x once per cycle.
SMEAR = r"""
module smear(input clk, input [3:0] a, output [15:0] o, output [15:0] ox);
    reg [15:0] x, y, r;
    always @* x = (16'h0001 << a) | (y << 1);
    always @* y = x;
    always @(posedge clk) r <= r + x;
    assign o = r;
    assign ox = x;
endmodule
"""


def smear_closed_form(stim, seed: int, sid: int, cycles: int) -> list:
    """Per-cycle output bytes (o then ox, 16 bits each, little-endian) of
    ``smear`` for stimulus ``sid``."""
    r = 0
    out = []
    for cy in range(cycles):
        a = stim(seed, sid, cy, 0) & 15
        x = (0xFFFF << a) & 0xFFFF
        r = (r + x) & 0xFFFF
        out.append(r.to_bytes(2, "little") + x.to_bytes(2, "little"))
    return out


def smear_passes(stim, seed: int, sid: int, cycles: int) -> list:
    """Settle passes of the clk = 0 eval in each cycle of ``smear``, counted
    by iterating the two always blocks from the previous fixed point until
    nothing changes (the pass that finds no change included)."""
    x = y = 0
    out = []
    for cy in range(cycles):
        a = stim(seed, sid, cy, 0) & 15
        n = 0
        while True:
            n += 1
            nx = ((1 << a) | (y << 1)) & 0xFFFF
            ny = nx
            if (nx, ny) == (x, y):
                break
            x, y = nx, ny
        out.append(n)
    return out


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
