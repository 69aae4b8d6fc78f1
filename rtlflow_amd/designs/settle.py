"""Settle-loop coverage designs (Verilator's UNOPTFLAT case): combinational
logic with a cycle through two always blocks.

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
