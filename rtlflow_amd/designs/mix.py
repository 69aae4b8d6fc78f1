"""A small design with ports of every shape the input stream and the output
capture lay out: narrow and single-bit ports, a 70-bit input (two stream
words) and a 100-bit output (two capture words), one clock."""

VERILOG = """
module mix(input clk, input [6:0] a, input [69:0] w, input b,
           output reg [99:0] r, output [12:0] s, output z);
  reg [15:0] cnt;
  always @(posedge clk) begin
    r <= {r[35:0], w[63:0]} ^ {30'd0, w};
    cnt <= cnt + {9'd0, a};
  end
  assign s = cnt[12:0] ^ {6'd0, a};
  assign z = b ^ cnt[0];
endmodule
"""
