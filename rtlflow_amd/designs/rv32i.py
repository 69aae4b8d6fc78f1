"""Synthetic design 3: a 5-stage in-order RV32I pipeline (project-written
Verilog) with a 64 KB instruction ROM and a 16 KB data RAM per stimulus, plus
the seeded random program generator and lognormal cycle budgets.

Pipeline specification (the C golden model in oracle/ follows this text):

* IF: ``d_ir <= imem[pc[15:2]]``, ``pc <= pc + 4``.
* ID: decode; register read with bypass from the instruction in WB.
* EX: ALU / address / branch with forwarding from MEM (non-load results) and
  WB; branches and jumps resolve here: when taken, the instructions in ID and
  IF are squashed and ``pc`` is redirected.
* MEM: data RAM word ``addr[13:2]``; byte / half lanes from ``addr[1:0]`` /
  ``addr[1]`` (accesses wrap in the 16 KB RAM and ignore misalignment);
  stores merge into the addressed word.
* WB: ``rf[rd] <= value`` (x0 never written).
* A load in EX whose rd equals rs1 or rs2 of the valid instruction in ID
  stalls IF and ID for one cycle (bubble into EX).
* Unknown opcodes execute as no-ops. All state starts at zero.

Outputs per cycle: ``pc_out`` = IF program counter, ``wb_out`` = the value the
WB-stage instruction writes (0 when it writes nothing).

Programs: 4,096 instructions drawn uniformly from ALU register/immediate ops,
LUI/AUIPC, loads, stores and forward branches (targets inside the program),
ending with the self-loop ``beq x0, x0, 0`` at index 4,095.
"""
from __future__ import annotations

import numpy as np

VERILOG = r"""
module rv32i(input clk, output [31:0] pc_out, output [31:0] wb_out);
    localparam OP = 7'b0110011, OPI = 7'b0010011, LUI = 7'b0110111, AUIPC = 7'b0010111,
               LOAD = 7'b0000011, STORE = 7'b0100011, BRANCH = 7'b1100011,
               JAL = 7'b1101111, JALR = 7'b1100111;

    reg [31:0] imem [0:16383];
    reg [31:0] dmem [0:4095];
    reg [31:0] rf [0:31];

    reg [31:0] pc;
    // IF/ID
    reg        d_v;
    reg [31:0] d_pc, d_ir;
    // ID/EX
    reg        e_v, e_wen, e_f7;
    reg [31:0] e_pc, e_a, e_b, e_imm;
    reg [4:0]  e_rd, e_rs1, e_rs2;
    reg [6:0]  e_op;
    reg [2:0]  e_f3;
    // EX/MEM
    reg        m_v, m_wen;
    reg [31:0] m_res, m_sd;
    reg [4:0]  m_rd;
    reg [6:0]  m_op;
    reg [2:0]  m_f3;
    // MEM/WB
    reg        w_wen;
    reg [31:0] w_res;
    reg [4:0]  w_rd;

    // ---------------------------------------------------------------- ID
    wire [6:0] d_op  = d_ir[6:0];
    wire [4:0] d_rd  = d_ir[11:7];
    wire [2:0] d_f3  = d_ir[14:12];
    wire [4:0] d_rs1 = d_ir[19:15];
    wire [4:0] d_rs2 = d_ir[24:20];
    reg  [31:0] d_imm;
    always @* begin
        case (d_op)
            STORE:       d_imm = {{20{d_ir[31]}}, d_ir[31:25], d_ir[11:7]};
            BRANCH:      d_imm = {{19{d_ir[31]}}, d_ir[31], d_ir[7], d_ir[30:25], d_ir[11:8], 1'b0};
            LUI, AUIPC:  d_imm = {d_ir[31:12], 12'b0};
            JAL:         d_imm = {{11{d_ir[31]}}, d_ir[31], d_ir[19:12], d_ir[20], d_ir[30:21], 1'b0};
            default:     d_imm = {{20{d_ir[31]}}, d_ir[31:20]};
        endcase
    end
    wire d_writes = (d_op == OP) || (d_op == OPI) || (d_op == LUI) || (d_op == AUIPC) ||
                    (d_op == LOAD) || (d_op == JAL) || (d_op == JALR);
    wire d_wen = d_v && d_writes && (d_rd != 5'd0);
    wire [31:0] d_a = (w_wen && (w_rd == d_rs1)) ? w_res : rf[d_rs1];
    wire [31:0] d_b = (w_wen && (w_rd == d_rs2)) ? w_res : rf[d_rs2];

    // ---------------------------------------------------------------- EX
    wire [31:0] x_a = (m_wen && (m_op != LOAD) && (m_rd == e_rs1)) ? m_res :
                      (w_wen && (w_rd == e_rs1)) ? w_res : e_a;
    wire [31:0] x_b = (m_wen && (m_op != LOAD) && (m_rd == e_rs2)) ? m_res :
                      (w_wen && (w_rd == e_rs2)) ? w_res : e_b;
    wire [31:0] x_op2 = (e_op == OP) ? x_b : e_imm;
    reg  [31:0] x_res;
    reg  [31:0] x_target;
    reg         x_taken;
    always @* begin
        x_res = 32'd0;
        x_taken = 1'b0;
        x_target = e_pc + e_imm;
        case (e_op)
            OP, OPI: begin
                case (e_f3)
                    3'd0: if ((e_op == OP) && e_f7) x_res = x_a - x_op2; else x_res = x_a + x_op2;
                    3'd1: x_res = x_a << x_op2[4:0];
                    3'd2: x_res = {31'd0, $signed(x_a) < $signed(x_op2)};
                    3'd3: x_res = {31'd0, x_a < x_op2};
                    3'd4: x_res = x_a ^ x_op2;
                    3'd5: if (e_f7) x_res = $signed(x_a) >>> x_op2[4:0]; else x_res = x_a >> x_op2[4:0];
                    3'd6: x_res = x_a | x_op2;
                    default: x_res = x_a & x_op2;
                endcase
            end
            LUI:   x_res = e_imm;
            AUIPC: x_res = e_pc + e_imm;
            LOAD, STORE: x_res = x_a + e_imm;
            BRANCH: begin
                case (e_f3)
                    3'd0: x_taken = x_a == x_b;
                    3'd1: x_taken = x_a != x_b;
                    3'd4: x_taken = $signed(x_a) < $signed(x_b);
                    3'd5: x_taken = $signed(x_a) >= $signed(x_b);
                    3'd6: x_taken = x_a < x_b;
                    3'd7: x_taken = x_a >= x_b;
                    default: x_taken = 1'b0;
                endcase
            end
            JAL: begin
                x_res = e_pc + 32'd4;
                x_taken = 1'b1;
            end
            JALR: begin
                x_res = e_pc + 32'd4;
                x_taken = 1'b1;
                x_target = (x_a + e_imm) & 32'hfffffffe;
            end
            default: x_res = 32'd0;
        endcase
        if (!e_v) x_taken = 1'b0;
    end
    wire load_use = e_v && (e_op == LOAD) && e_wen && d_v && ((e_rd == d_rs1) || (e_rd == d_rs2));

    // --------------------------------------------------------------- MEM
    wire [31:0] m_word = dmem[m_res[13:2]];
    wire [7:0]  m_byte = m_word[{m_res[1:0], 3'b000} +: 8];
    wire [15:0] m_half = m_res[1] ? m_word[31:16] : m_word[15:0];
    reg  [31:0] m_load;
    reg  [31:0] m_merged;
    always @* begin
        case (m_f3)
            3'd0: m_load = {{24{m_byte[7]}}, m_byte};
            3'd1: m_load = {{16{m_half[15]}}, m_half};
            3'd4: m_load = {24'd0, m_byte};
            3'd5: m_load = {16'd0, m_half};
            default: m_load = m_word;
        endcase
        m_merged = m_word;
        case (m_f3[1:0])
            2'd0: m_merged[{m_res[1:0], 3'b000} +: 8] = m_sd[7:0];
            2'd1: if (m_res[1]) m_merged[31:16] = m_sd[15:0]; else m_merged[15:0] = m_sd[15:0];
            default: m_merged = m_sd;
        endcase
    end

    // ------------------------------------------------------------ clocked
    always @(posedge clk) begin
        // WB
        if (w_wen) rf[w_rd] <= w_res;
        // MEM -> WB
        w_wen <= m_wen;
        w_rd  <= m_rd;
        w_res <= (m_op == LOAD) ? m_load : m_res;
        if (m_v && (m_op == STORE)) dmem[m_res[13:2]] <= m_merged;
        // EX -> MEM
        m_v   <= e_v;
        m_wen <= e_v && e_wen;
        m_res <= x_res;
        m_sd  <= x_b;
        m_rd  <= e_rd;
        m_op  <= e_op;
        m_f3  <= e_f3;
        // ID -> EX
        if (x_taken || load_use) begin
            e_v <= 1'b0;
            e_wen <= 1'b0;
            e_op <= 7'd0;
        end else begin
            e_v   <= d_v;
            e_wen <= d_wen;
            e_pc  <= d_pc;
            e_a   <= d_a;
            e_b   <= d_b;
            e_imm <= d_imm;
            e_rd  <= d_rd;
            e_rs1 <= d_rs1;
            e_rs2 <= d_rs2;
            e_op  <= d_op;
            e_f3  <= d_f3;
            e_f7  <= d_ir[30];
        end
        // IF -> ID
        if (x_taken) begin
            d_v <= 1'b0;
        end else if (!load_use) begin
            d_v  <= 1'b1;
            d_pc <= pc;
            d_ir <= imem[pc[15:2]];
        end
        // PC
        if (x_taken) pc <= x_target;
        else if (!load_use) pc <= pc + 32'd4;
    end

    assign pc_out = pc;
    assign wb_out = w_wen ? w_res : 32'd0;
endmodule
"""

PROGRAM_WORDS = 4096
ROM_WORDS = 16384
HALT = 0x00000063  # beq x0, x0, 0

_M64 = (1 << 64) - 1
_G = 0x9E3779B97F4A7C15


def verilog() -> str:
    return VERILOG


def _mix(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finalizer, vectorized over uint64 arrays."""
    z = z + np.uint64(_G)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _rand(seed: int, sids: np.ndarray, n: int, stream: int) -> np.ndarray:
    """[len(sids), n] uint64 draws: splitmix(seed ^ s*G ^ (i << 8) ^ stream)."""
    with np.errstate(over="ignore"):
        s = sids.astype(np.uint64)[:, None] * np.uint64(_G)
        i = (np.arange(n, dtype=np.uint64) << np.uint64(8))[None, :]
        return _mix(np.uint64(seed) ^ s ^ i ^ np.uint64(stream))


def programs(seed: int, sids, words: int = PROGRAM_WORDS) -> np.ndarray:
    """Random valid programs, one per stimulus id: uint32 [len(sids), words]."""
    sids = np.asarray(sids, dtype=np.uint64)
    with np.errstate(over="ignore"):
        r0 = _rand(seed, sids, words, 0x11)
        r1 = _rand(seed, sids, words, 0x22)
    cls = (r0 % np.uint64(7)).astype(np.int64)
    rd = ((r0 >> np.uint64(8)) & np.uint64(31)).astype(np.uint32)
    rs1 = ((r0 >> np.uint64(16)) & np.uint64(31)).astype(np.uint32)
    rs2 = ((r0 >> np.uint64(24)) & np.uint64(31)).astype(np.uint32)
    sel = ((r0 >> np.uint64(32)) & np.uint64(0xffff)).astype(np.uint32)
    imm12 = (r1 & np.uint64(0xfff)).astype(np.uint32)
    imm20 = ((r1 >> np.uint64(12)) & np.uint64(0xfffff)).astype(np.uint32)
    boff = (((r1 >> np.uint64(40)) % np.uint64(16)).astype(np.uint32) + 1) * 4  # 4..64 bytes ahead

    out = np.zeros(rd.shape, np.uint32)
    # R-type ALU: (f3, f7)
    r_ops = np.array([(0, 0), (0, 0x20), (1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (5, 0x20), (6, 0), (7, 0)],
                     np.uint32)
    k = sel % len(r_ops)
    rtype = (r_ops[k, 1] << 25) | (rs2 << 20) | (rs1 << 15) | (r_ops[k, 0] << 12) | (rd << 7) | 0x33
    # I-type ALU
    i_f3 = np.array([0, 2, 3, 4, 6, 7, 1, 5, 5], np.uint32)
    ki = sel % 9
    f3 = i_f3[ki]
    shamt = imm12 & 31
    imm_i = np.where(ki == 6, shamt, np.where(ki == 7, shamt, np.where(ki == 8, shamt | 0x400, imm12)))
    itype = (imm_i << 20) | (rs1 << 15) | (f3 << 12) | (rd << 7) | 0x13
    # U-type
    utype = (imm20 << 12) | (rd << 7) | np.where(sel & 1, 0x17, 0x37).astype(np.uint32)
    # loads
    l_f3 = np.array([0, 1, 2, 4, 5], np.uint32)[sel % 5]
    ltype = (imm12 << 20) | (rs1 << 15) | (l_f3 << 12) | (rd << 7) | 0x03
    # stores
    s_f3 = np.array([0, 1, 2], np.uint32)[sel % 3]
    stype = (((imm12 >> 5) & 0x7f) << 25) | (rs2 << 20) | (rs1 << 15) | (s_f3 << 12) | \
        ((imm12 & 31) << 7) | 0x23
    # forward branches, target kept inside the program
    b_f3 = np.array([0, 1, 4, 5, 6, 7], np.uint32)[sel % 6]
    idx = np.arange(words, dtype=np.uint32)[None, :]
    limit = (np.uint32(words - 1) - idx) * 4
    off = np.minimum(boff, np.maximum(limit, 4))
    btype = (((off >> 12) & 1) << 31) | (((off >> 5) & 0x3f) << 25) | (rs2 << 20) | (rs1 << 15) | \
        (b_f3 << 12) | (((off >> 1) & 0xf) << 8) | (((off >> 11) & 1) << 7) | 0x63
    choices = [rtype, itype, itype, utype, ltype, stype, btype]
    for c, arr in enumerate(choices):
        out = np.where(cls == c, arr, out)
    out[:, words - 1] = HALT
    return out.astype(np.uint32)


def budgets(seed: int, n: int, median: int = 5000, sigma: float = 1.0, lo: int = 1000,
            hi: int = 100000) -> np.ndarray:
    """Lognormal cycle budgets (median ``median``), clipped and rounded; from
    a NumPy generator seeded with ``seed`` so CPU and GPU get the same array."""
    rng = np.random.default_rng(seed & 0xffffffff)
    x = np.exp(np.log(median) + sigma * rng.standard_normal(n))
    return np.clip(np.rint(x), lo, hi).astype(np.uint32)
