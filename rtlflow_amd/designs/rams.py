"""Memory-port coverage design: synchronous reads into registers, guarded
writes under if / case, two write ports on one memory in one clocked block
(the later one wins at a shared address), reads of the element written at
the same edge (old data), 8-, 6- and 64-bit memories of depths that are not
powers of two (out-of-range addresses read 0 and do not write)."""

VERILOG = r"""
module rams(input clk, input [6:0] a, input [6:0] b, input [7:0] d, input [1:0] op, input [63:0] w,
            output [7:0] o, output [63:0] ow, output [5:0] oc);
    reg [7:0] m8 [0:99];
    reg [63:0] m64 [0:9];
    reg [5:0] mc [0:4];
    reg [7:0] r1, r2;
    reg [63:0] r3;
    reg [5:0] r4;
    reg [7:0] cnt;
    always @(posedge clk) begin
        cnt <= cnt + 8'd1;
        r1 <= m8[a];
        r2 <= m8[b ^ cnt[6:0]];
        if (op[0]) begin
            m8[a] <= d;
            if (op[1]) m8[b] <= d ^ cnt;
        end else
            m8[cnt[6:0]] <= r1 + r2;
        case (op)
            2'd0: m64[a[3:0]] <= w;
            2'd2: m64[b[3:0]] <= {w[31:0], w[63:32]};
            default: ;
        endcase
        r3 <= m64[a[3:0] ^ b[3:0]];
        mc[cnt[2:0]] <= d[5:0] ^ r4;
        r4 <= mc[a[2:0]];
    end
    assign o = r1 ^ r2;
    assign ow = r3;
    assign oc = r4;
endmodule
"""


def verilog() -> str:
    return VERILOG
