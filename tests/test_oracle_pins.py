"""Pins of the per-design golden models (oracle/oracle_designs.c) by means
independent of their own per-cycle formulas:

* the IR interpreter (oracle/ir_interp.py, itself pinned by known-answer
  expressions in test_semantics.py) runs the designs' Verilog -- the text the
  GPU path compiles -- and must give the C models' hashes (ALU, systolic,
  RV32I with program images and budgets);
* the systolic model against a closed form: each accumulator is the sum of
  its skewed int8 products, the chain is a window of the last 512 input
  bytes;
* the RV32I instruction-set simulator and pipeline model against
  hand-assembled programs whose register and memory results are computed by
  hand (SLTIU / SLTI / SRA / SRAI, LUI / AUIPC, LB / LH / LBU / LHU sign and
  zero extension, SB / SH merges, backward branches).
"""
import ctypes as C

import numpy as np

import ir_interp
import oracle
from rtlflow_amd.designs import alu, registry, rv32i, systolic
from rtlflow_amd.verilog import elaborate

M32 = 0xFFFFFFFF


# ----------------------------------------------------------------- interpreter
def test_interpreter_on_alu_verilog_matches_alu_model():
    seed = registry.SEED_CFG[1]
    want = list(oracle.alu_run(seed, 4, 300, stim0=5))
    assert ir_interp.protocol(elaborate(alu.verilog()), seed, 4, 300, stim0=5) == want


def test_interpreter_on_systolic_verilog_matches_systolic_model():
    seed = registry.SEED_CFG[4]
    want = list(oracle.systolic_run(seed, 2, 40, stim0=1))
    assert ir_interp.protocol(elaborate(systolic.verilog()), seed, 2, 40, stim0=1) == want


def test_interpreter_on_rv32i_verilog_matches_rv32i_model():
    seed = registry.SEED_CFG[3]
    progs = rv32i.programs(seed, np.arange(3))
    budgets = [1200, 0, 700]
    want = list(oracle.rv32i_run(progs, seed, budgets))
    got = ir_interp.protocol(elaborate(rv32i.verilog()), seed, 3, 0, images={"imem": progs},
                             budgets=budgets)
    assert got == want


# ------------------------------------------------------- systolic closed form
def _systolic_closed_hash(seed, s, cycles, n=16, window=512):
    """Hash of the systolic design computed from closed forms. Operand a of
    PE (i, j) at the edge of cycle t is row byte i of cycle t - 1 - j, and
    b is column byte j of cycle t - 1 - i (zero before cycle 0), so after
    cycle c, acc(i, j) = sum over t <= c of sx(row_i[t-1-j]) * sx(col_j[t-1-i]).
    The chain before the edge of cycle t holds the bytes row_0 ^ col_0 of
    cycles t-1, t-2, ... (byte k = cycle t-1-k): cmp compares its upper 256
    bytes with its lower 256 as integers, par is the parity of its bits."""
    rows = np.zeros((cycles, n), np.int64)
    cols = np.zeros((cycles, n), np.int64)
    for c in range(cycles):
        r = oracle.stim(seed, s, c, 0) | (oracle.stim(seed, s, c, 1) << 64)
        q = oracle.stim(seed, s, c, 8) | (oracle.stim(seed, s, c, 9) << 64)
        rows[c] = [(r >> (8 * i)) & 0xFF for i in range(n)]
        cols[c] = [(q >> (8 * j)) & 0xFF for j in range(n)]
    sx = lambda v: np.where(v >= 128, v - 256, v)
    R, Q = sx(rows), sx(cols)
    bytes_in = (rows[:, 0] ^ cols[:, 0]).astype(np.int64)
    acc = np.zeros((n, n), np.int64)
    h = 0xCBF29CE484222325
    for c in range(cycles):
        # add the products that enter at the edge of cycle c
        for i in range(n):
            for j in range(n):
                ta, tb = c - 1 - j, c - 1 - i
                if ta >= 0 and tb >= 0:
                    acc[i, j] += R[ta, i] * Q[tb, j]
        x = 0
        for v in acc.flatten():
            x ^= int(v) & M32
        win = [int(bytes_in[c - 1 - k]) if c - 1 - k >= 0 else 0 for k in range(window)]
        lo = int.from_bytes(bytes(win[:window // 2]), "little")
        hi = int.from_bytes(bytes(win[window // 2:]), "little")
        cmp_ = int(hi < lo)
        par = sum(bin(b).count("1") for b in win) & 1
        h = oracle.fnv(x.to_bytes(4, "little"), h)
        h = oracle.fnv(bytes([(cmp_ << 1) | par]), h)
    return h


def test_systolic_model_matches_closed_form():
    """600 cycles: the 512-byte chain window fills and shifts."""
    seed = registry.SEED_CFG[4]
    got = oracle.systolic_run(seed, 2, 600, stim0=7)
    for k in range(2):
        assert int(got[k]) == _systolic_closed_hash(seed, 7 + k, 600)


def test_systolic_closed_form_accumulators_are_skewed_products():
    """The closed form on a hand case: row byte 2 of cycle 0 is 3 and column
    byte 5 of cycle 3 is -2. The row byte sits in PE (2, j) at the edge of
    cycle j + 1, the column byte in PE (i, 5) at the edge of cycle i + 4, so
    they meet once, in PE (2, 5) at cycle 6, and no other product is
    nonzero."""
    n = 16
    R = np.zeros((10, n), np.int64)
    Q = np.zeros((10, n), np.int64)
    R[0, 2] = 3
    Q[3, 5] = -2
    acc = np.zeros((n, n), np.int64)
    for c in range(10):
        for i in range(n):
            for j in range(n):
                ta, tb = c - 1 - j, c - 1 - i
                if ta >= 0 and tb >= 0:
                    acc[i, j] += R[ta, i] * Q[tb, j]
    assert acc[2, 5] == -6 and np.count_nonzero(acc) == 1


# --------------------------------------------------------- RV32I known answers
def _r(op, rd, f3, rs1, rs2, f7=0):
    return (f7 << 25) | (rs2 << 20) | (rs1 << 15) | (f3 << 12) | (rd << 7) | op


def _i(op, rd, f3, rs1, imm):
    return ((imm & 0xFFF) << 20) | (rs1 << 15) | (f3 << 12) | (rd << 7) | op


def _s(rs1, rs2, f3, imm):
    imm &= 0xFFF
    return ((imm >> 5) << 25) | (rs2 << 20) | (rs1 << 15) | (f3 << 12) | ((imm & 31) << 7) | 0x23


def _b(rs1, rs2, f3, off):
    off &= 0x1FFF
    return (((off >> 12) & 1) << 31) | (((off >> 5) & 0x3F) << 25) | (rs2 << 20) | (rs1 << 15) | \
        (f3 << 12) | (((off >> 1) & 0xF) << 8) | (((off >> 11) & 1) << 7) | 0x63


def _u(op, rd, imm20):
    return ((imm20 & 0xFFFFF) << 12) | (rd << 7) | op


addi = lambda rd, rs, imm: _i(0x13, rd, 0, rs, imm)
slti = lambda rd, rs, imm: _i(0x13, rd, 2, rs, imm)
sltiu = lambda rd, rs, imm: _i(0x13, rd, 3, rs, imm)
xori = lambda rd, rs, imm: _i(0x13, rd, 4, rs, imm)
ori = lambda rd, rs, imm: _i(0x13, rd, 6, rs, imm)
andi = lambda rd, rs, imm: _i(0x13, rd, 7, rs, imm)
slli = lambda rd, rs, sh: _i(0x13, rd, 1, rs, sh)
srli = lambda rd, rs, sh: _i(0x13, rd, 5, rs, sh)
srai = lambda rd, rs, sh: _i(0x13, rd, 5, rs, 0x400 | sh)
add = lambda rd, a, b: _r(0x33, rd, 0, a, b)
sub = lambda rd, a, b: _r(0x33, rd, 0, a, b, 0x20)
slt = lambda rd, a, b: _r(0x33, rd, 2, a, b)
sltu = lambda rd, a, b: _r(0x33, rd, 3, a, b)
sra = lambda rd, a, b: _r(0x33, rd, 5, a, b, 0x20)
lui = lambda rd, imm20: _u(0x37, rd, imm20)
auipc = lambda rd, imm20: _u(0x17, rd, imm20)
lb = lambda rd, rs, off: _i(0x03, rd, 0, rs, off)
lh = lambda rd, rs, off: _i(0x03, rd, 1, rs, off)
lw = lambda rd, rs, off: _i(0x03, rd, 2, rs, off)
lbu = lambda rd, rs, off: _i(0x03, rd, 4, rs, off)
lhu = lambda rd, rs, off: _i(0x03, rd, 5, rs, off)
sb = lambda rs2, rs1, off: _s(rs1, rs2, 0, off)
sh = lambda rs2, rs1, off: _s(rs1, rs2, 1, off)
sw = lambda rs2, rs1, off: _s(rs1, rs2, 2, off)
beq = lambda a, b, off: _b(a, b, 0, off)
bne = lambda a, b, off: _b(a, b, 1, off)
blt = lambda a, b, off: _b(a, b, 4, off)
bge = lambda a, b, off: _b(a, b, 5, off)
bltu = lambda a, b, off: _b(a, b, 6, off)
bgeu = lambda a, b, off: _b(a, b, 7, off)
HALT = beq(0, 0, 0)

PROG_ALU = [
    addi(1, 0, -5),          # x1 = 0xfffffffb
    sltiu(2, 1, 3),          # 0xfffffffb <u 3: 0
    sltiu(3, 0, 1),          # 0 <u 1: 1
    slti(4, 1, 3),           # -5 < 3: 1
    srai(5, 1, 1),           # -5 >>> 1 = -3
    srli(6, 1, 28),          # 0xf
    addi(7, 0, 4),
    sra(8, 1, 7),            # -5 >>> 4 = -1
    lui(9, 0x80000),         # 0x80000000
    sra(10, 9, 7),           # 0xf8000000
    sltu(11, 0, 1),          # 0 <u 0xfffffffb: 1
    slt(12, 1, 0),           # -5 < 0: 1
    sub(13, 0, 1),           # 5
    xori(14, 1, -1),         # ~x1 = 4
    andi(15, 1, 0x7F0),      # 0x7f0
    ori(16, 0, -2048),       # 0xfffff800
    slli(17, 13, 30),        # 5 << 30 mod 2^32 = 0x40000000
    auipc(18, 1),            # pc 68 + 0x1000
    HALT,
]
WANT_ALU = {1: 0xFFFFFFFB, 2: 0, 3: 1, 4: 1, 5: 0xFFFFFFFD, 6: 0xF, 7: 4, 8: 0xFFFFFFFF,
            9: 0x80000000, 10: 0xF8000000, 11: 1, 12: 1, 13: 5, 14: 4, 15: 0x7F0,
            16: 0xFFFFF800, 17: 0x40000000, 18: 0x1044}

PROG_MEM = [
    lui(1, 0x12345),         # 0x12345000
    addi(1, 1, 0x678),       # 0x12345678
    addi(2, 0, 0x80),        # base address 128 = word 32
    sw(1, 2, 0),             # word 32 = 0x12345678
    addi(3, 0, -1),
    sb(3, 2, 5),             # byte 133 = 0xff: word 33 = 0x0000ff00
    addi(4, 0, 0x7FF),
    sh(4, 2, 6),             # half 134 = 0x07ff: word 33 = 0x07ffff00
    lb(5, 2, 0),             # byte 128 = 0x78
    lb(6, 2, 5),             # 0xff -> -1
    lbu(7, 2, 5),            # 0xff
    lh(8, 2, 2),             # half 130 = 0x1234
    lh(9, 2, 4),             # half 132 = 0xff00 -> 0xffffff00
    lhu(10, 2, 4),           # 0xff00
    lw(11, 2, 4),            # 0x07ffff00
    lb(12, 2, 3),            # byte 131 = 0x12
    sb(1, 2, 11),            # byte 139 = 0x78: word 34 = 0x78000000
    sh(1, 2, 12),            # half 140 = 0x5678: word 35 = 0x00005678
    HALT,
]
WANT_MEM_RF = {1: 0x12345678, 2: 0x80, 3: 0xFFFFFFFF, 4: 0x7FF, 5: 0x78, 6: 0xFFFFFFFF, 7: 0xFF,
               8: 0x1234, 9: 0xFFFFFF00, 10: 0xFF00, 11: 0x07FFFF00, 12: 0x12}
WANT_MEM_DM = {32: 0x12345678, 33: 0x07FFFF00, 34: 0x78000000, 35: 0x00005678}

PROG_BR = [
    addi(1, 0, 0),           # 0: i = 0
    addi(2, 0, 10),          # 4: n = 10
    addi(3, 0, 0),           # 8: sum = 0
    add(3, 3, 1),            # 12: loop: sum += i
    addi(1, 1, 1),           # 16: i += 1
    blt(1, 2, -8),           # 20: back to 12 while i < n
    addi(4, 0, -3),          # 24
    bge(4, 0, 8),            # 28: -3 >= 0 is false
    addi(5, 0, 1),           # 32: x5 = 1
    bltu(4, 2, 12),          # 36: 0xfffffffd <u 10 is false
    addi(6, 0, 2),           # 40: x6 = 2
    bne(6, 0, 8),            # 44: taken, to 52
    addi(7, 0, 99),          # 48: skipped
    bgeu(4, 2, 8),           # 52: taken, to 60
    addi(8, 0, 77),          # 56: skipped
    addi(9, 0, 5),           # 60: countdown from 5
    addi(10, 10, 3),         # 64: x10 += 3
    addi(9, 9, -1),          # 68
    bne(9, 0, -8),           # 72: back to 64 while x9 != 0
    HALT,                    # 76
]
WANT_BR = {1: 10, 2: 10, 3: 45, 4: 0xFFFFFFFD, 5: 1, 6: 2, 7: 0, 8: 0, 9: 0, 10: 15}


def _image(prog):
    img = np.zeros(4096, np.uint32)
    img[:len(prog)] = [p & M32 for p in prog]
    return img


def _iss(prog):
    L = oracle.lib()
    L.or_rv32i_iss.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    img = _image(prog)
    rf, dm = np.zeros(32, np.uint32), np.zeros(4096, np.uint32)
    n = L.or_rv32i_iss(img.ctypes.data, 4096, 4 * (len(prog) - 1), 10000, rf.ctypes.data, dm.ctypes.data)
    return n, rf, dm


def _pipeline(prog, cycles=400):
    L = oracle.lib()
    L.or_rv32i_state.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    img = _image(prog)
    rf, dm = np.zeros(32, np.uint32), np.zeros(4096, np.uint32)
    L.or_rv32i_state(img.ctypes.data, 4096, cycles, rf.ctypes.data, dm.ctypes.data)
    return rf, dm


def _check_rf(rf, want):
    for r in range(32):
        assert int(rf[r]) == want.get(r, 0), (r, hex(int(rf[r])), hex(want.get(r, 0)))


def test_rv32i_hand_program_alu_ops():
    n, rf, dm = _iss(PROG_ALU)
    assert n == len(PROG_ALU) - 1
    _check_rf(rf, WANT_ALU)
    rf2, dm2 = _pipeline(PROG_ALU)
    _check_rf(rf2, WANT_ALU)
    assert not dm.any() and not dm2.any()


def test_rv32i_hand_program_loads_and_stores():
    n, rf, dm = _iss(PROG_MEM)
    _check_rf(rf, WANT_MEM_RF)
    rf2, dm2 = _pipeline(PROG_MEM)
    _check_rf(rf2, WANT_MEM_RF)
    for d in (dm, dm2):
        nz = {int(k): int(d[k]) for k in np.nonzero(d)[0]}
        assert nz == WANT_MEM_DM


def test_rv32i_hand_program_backward_branches():
    n, rf, dm = _iss(PROG_BR)
    # 3 set-up + 10 x 3 loop + 2 + 1 + 1 + 1 + 1 (bne) + 1 (bgeu) + 1 + 5 x 3
    assert n == 3 + 30 + 2 + 1 + 1 + 1 + 1 + 1 + 1 + 15
    _check_rf(rf, WANT_BR)
    rf2, _ = _pipeline(PROG_BR)
    _check_rf(rf2, WANT_BR)


def test_rv32i_hand_programs_through_verilog():
    """The same programs in the RV32I Verilog (IR interpreter) hash as the
    pipeline model does, so the hand-checked architectural results pin the
    design the GPU runs as well."""
    seed = registry.SEED_CFG[3]
    progs = np.stack([_image(p) for p in (PROG_ALU, PROG_MEM, PROG_BR)])
    budgets = [120, 120, 300]
    want = list(oracle.rv32i_run(progs, seed, budgets))
    got = ir_interp.protocol(elaborate(rv32i.verilog()), seed, 3, 0, images={"imem": progs},
                             budgets=budgets)
    assert got == want
