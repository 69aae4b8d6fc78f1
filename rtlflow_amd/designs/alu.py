"""Synthetic design 1: a 4-stage pipelined 32-bit ALU with a 32 x 32-bit
register file (project-written Verilog).

Per cycle the testbench drives ``op`` (3 bits), ``rs1``/``rs2``/``rd`` (5 bits)
and ``imm`` (32 bits). Stage 1 latches the instruction, stage 2 reads the
register file (``b`` operand = rf[rs2] ^ imm), stage 3 executes one of
add/sub/and/or/xor/shl/shr/mul, stage 4 writes the result back to rf[rd] and
drives ``result``. There is no forwarding: a read in stage 2 sees the register
file before the same edge's write-back (non-blocking semantics).
"""

VERILOG = r"""
module alu(
    input clk,
    input [2:0] op,
    input [4:0] rs1,
    input [4:0] rs2,
    input [4:0] rd,
    input [31:0] imm,
    output [31:0] result
);
    reg [31:0] rf [0:31];
    reg [2:0]  s1_op;
    reg [4:0]  s1_rs1, s1_rs2, s1_rd;
    reg [31:0] s1_imm;
    reg [2:0]  s2_op;
    reg [4:0]  s2_rd;
    reg [31:0] s2_a, s2_b;
    reg [4:0]  s3_rd;
    reg [31:0] s3_res;
    reg [31:0] s4_res;

    always @(posedge clk) begin
        // stage 1: latch the instruction
        s1_op  <= op;
        s1_rs1 <= rs1;
        s1_rs2 <= rs2;
        s1_rd  <= rd;
        s1_imm <= imm;
        // stage 2: register read
        s2_op <= s1_op;
        s2_rd <= s1_rd;
        s2_a  <= rf[s1_rs1];
        s2_b  <= rf[s1_rs2] ^ s1_imm;
        // stage 3: execute
        s3_rd <= s2_rd;
        case (s2_op)
            3'd0: s3_res <= s2_a + s2_b;
            3'd1: s3_res <= s2_a - s2_b;
            3'd2: s3_res <= s2_a & s2_b;
            3'd3: s3_res <= s2_a | s2_b;
            3'd4: s3_res <= s2_a ^ s2_b;
            3'd5: s3_res <= s2_a << s2_b[4:0];
            3'd6: s3_res <= s2_a >> s2_b[4:0];
            default: s3_res <= s2_a * s2_b;
        endcase
        // stage 4: write back
        rf[s3_rd] <= s3_res;
        s4_res <= s3_res;
    end

    assign result = s4_res;
endmodule
"""


def verilog() -> str:
    return VERILOG
