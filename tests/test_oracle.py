"""Pins of the CPU oracle's primitives against published known answers and
independent restatements (the oracle is only trusted after these pass)."""
import numpy as np
import pytest

import oracle
from rtlflow_amd.designs.netlist import AND, MUX, NAND, OR, XOR, generate

M64 = (1 << 64) - 1


def test_fnv1a64_known_answers():
    # published FNV-1a 64 test vectors (Fowler/Noll/Vo)
    assert oracle.fnv(b"") == 0xcbf29ce484222325
    assert oracle.fnv(b"a") == 0xaf63dc4c8601ec8c
    assert oracle.fnv(b"foobar") == 0x85944171f73967e8


def test_splitmix64_known_answers():
    # SplitMix64 seeded with 0: first outputs of the reference generator
    assert oracle.splitmix64(0) == 0xe220a8397b1dcdaf
    assert oracle.splitmix64(0x9E3779B97F4A7C15) == 0x6e789e6aa1b965f4


def _py_splitmix(x):
    z = (x + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_stimulus_counter_definition():
    rng = np.random.default_rng(1)
    for _ in range(200):
        seed, s, c, k = (int(x) for x in rng.integers(0, 2**63, 4))
        s %= 1 << 20
        c %= 1 << 24
        k %= 256
        want = _py_splitmix(seed ^ ((s * 0x9E3779B97F4A7C15) & M64) ^ ((c << 8) & M64) ^ k)
        assert oracle.stim(seed, s, c, k) == want


def _py_netlist(table, seed, s, cycles):
    """Brute-force restatement of the netlist semantics in Python: primary
    input bits [64j, 64j + 64) from draw k = j; flip-flops and each RAM's
    registered read port update at the rising edge from the clk = 0 values,
    the read port taking the word from before the same edge's write."""
    n_first = table.n_pi + table.n_ff + table.n_ram_bits
    v = [0] * (n_first + table.n_gates)
    q = [0] * table.n_ff
    n_rams, w = len(table.ram_we), table.ram_width
    mem = [[0] * table.ram_depth for _ in range(n_rams)]
    rd = [0] * n_rams
    h = 0xcbf29ce484222325

    def comb(pi):
        for i in range(table.n_pi):
            v[i] = (pi >> i) & 1
        for i in range(table.n_ff):
            v[table.n_pi + i] = q[i]
        for r in range(n_rams):
            for k in range(w):
                v[table.n_pi + table.n_ff + r * w + k] = (rd[r] >> k) & 1
        for g in range(table.n_gates):
            a, b, c = (v[int(x)] for x in table.gate_in[g])
            t = int(table.gate_type[g])
            v[n_first + g] = {AND: a & b, OR: a | b, XOR: a ^ b, NAND: 1 - (a & b),
                              MUX: b if a else c}[t]

    for cyc in range(cycles):
        pi = 0
        for j in range((table.n_pi + 63) // 64):
            pi |= _py_splitmix(seed ^ ((s * 0x9E3779B97F4A7C15) & M64) ^ (cyc << 8) ^ j) << (64 * j)
        pi &= (1 << table.n_pi) - 1
        comb(pi)
        d = [v[int(x)] for x in table.ff_d]
        for r in range(n_rams):
            addr = sum(v[int(x)] << k for k, x in enumerate(table.ram_addr[r]))
            data = sum(v[int(x)] << k for k, x in enumerate(table.ram_wdata[r]))
            old = mem[r][addr] if addr < table.ram_depth else 0
            if v[int(table.ram_we[r])] and addr < table.ram_depth:
                mem[r][addr] = data
            rd[r] = old
        q[:] = d
        comb(pi)
        po = sum(v[int(x)] << k for k, x in enumerate(table.po))
        nb = 1 if len(table.po) <= 8 else 2 if len(table.po) <= 16 else 4 if len(table.po) <= 32 else 8
        h = oracle.fnv(po.to_bytes(nb, "little"), h)
    return h


def test_netlist_oracle_matches_bruteforce():
    _, table = generate(n_gates=120, n_ff=12, n_pi=10, n_po=9, depth=5, seed=3)
    seed = 0x52544C4600000002
    got = oracle.netlist_run(table.to_bytes(), seed, nstim=5, cycles=7)
    for s in range(5):
        assert int(got[s]) == _py_netlist(table, seed, s, 7)


def test_netlist_oracle_rams_and_wide_inputs_match_bruteforce():
    """Two input draws (70 primary inputs) and three synchronous RAMs of a
    depth that is not a power of two (out-of-range addresses read 0 and do
    not write)."""
    _, table = generate(n_gates=160, n_ff=10, n_pi=70, n_po=12, depth=5, seed=21, n_rams=3,
                        ram_depth=6, ram_width=5)
    seed = 0x52544C4600000005
    got = oracle.netlist_run(table.to_bytes(), seed, nstim=4, cycles=40)
    for s in range(4):
        assert int(got[s]) == _py_netlist(table, seed, s, 40)


def test_netlist_oracle_budgets_and_stim0():
    _, table = generate(n_gates=80, n_ff=8, n_pi=8, n_po=8, depth=4, seed=9)
    tb = table.to_bytes()
    full = oracle.netlist_run(tb, 11, nstim=6, cycles=5)
    part = oracle.netlist_run(tb, 11, nstim=3, cycles=5, stim0=3)
    assert (full[3:] == part).all()
    b = oracle.netlist_run(tb, 11, nstim=6, budgets=[5, 0, 5, 1, 5, 2])
    assert b[0] == full[0] and b[1] == 0xcbf29ce484222325


def _py_alu(seed, s, cycles):
    """Independent restatement of the ALU pipeline in Python."""
    rf = [0] * 32
    s1 = [0, 0, 0, 0, 0]
    s2 = [0, 0, 0, 0]
    s3 = [0, 0]
    s4 = 0
    h = 0xcbf29ce484222325
    M = 0xffffffff
    for c in range(cycles):
        op, rs1, rs2, rd, imm = (_py_splitmix(seed ^ ((s * 0x9E3779B97F4A7C15) & M64) ^ (c << 8) ^ (p << 3))
                                 for p in range(5))
        op, rs1, rs2, rd, imm = op & 7, rs1 & 31, rs2 & 31, rd & 31, imm & M
        o, a, b = s2[0], s2[2], s2[3]
        res = [(a + b) & M, (a - b) & M, a & b, a | b, a ^ b, (a << (b & 31)) & M, a >> (b & 31),
               (a * b) & M][o]
        n_rf = list(rf)
        n_rf[s3[0]] = s3[1]
        n_s4 = s3[1]
        n_s3 = [s2[1], res]
        n_s2 = [s1[0], s1[3], rf[s1[1]], rf[s1[2]] ^ s1[4]]
        n_s1 = [op, rs1, rs2, rd, imm]
        rf, s1, s2, s3, s4 = n_rf, n_s1, n_s2, n_s3, n_s4
        h = oracle.fnv(s4.to_bytes(4, "little"), h)
    return h


def test_alu_oracle_matches_python_restatement():
    seed = 0x52544C4600000001
    got = oracle.alu_run(seed, 4, 300)
    for s in range(4):
        assert int(got[s]) == _py_alu(seed, s, 300)


def _py_systolic(seed, s, cycles, n=16, chain_bits=4096):
    """Independent restatement of the systolic array + chain with Python ints."""
    a = [[0] * n for _ in range(n)]
    b = [[0] * n for _ in range(n)]
    acc = [[0] * n for _ in range(n)]
    chain = 0
    cmp_ = par = 0
    h = 0xcbf29ce484222325
    sx = lambda v: v - 256 if v >= 128 else v
    half = chain_bits // 2
    for c in range(cycles):
        r = [_py_splitmix(seed ^ ((s * 0x9E3779B97F4A7C15) & M64) ^ (c << 8) ^ k) for k in (0, 1, 8, 9)]
        row = (r[0] | (r[1] << 64))
        col = (r[2] | (r[3] << 64))
        rowb = [(row >> (8 * i)) & 0xff for i in range(n)]
        colb = [(col >> (8 * j)) & 0xff for j in range(n)]
        na = [[rowb[i] if j == 0 else a[i][j - 1] for j in range(n)] for i in range(n)]
        nb = [[colb[j] if i == 0 else b[i - 1][j] for j in range(n)] for i in range(n)]
        nacc = [[(acc[i][j] + sx(a[i][j]) * sx(b[i][j])) & 0xffffffff for j in range(n)] for i in range(n)]
        ncmp = int((chain >> half) < (chain & ((1 << half) - 1)))
        npar = bin(chain).count("1") & 1
        chain = ((chain << 8) | (rowb[0] ^ colb[0])) & ((1 << chain_bits) - 1)
        a, b, acc, cmp_, par = na, nb, nacc, ncmp, npar
        x = 0
        for i in range(n):
            for j in range(n):
                x ^= acc[i][j]
        h = oracle.fnv(x.to_bytes(4, "little"), h)
        h = oracle.fnv(bytes([(cmp_ << 1) | par]), h)
    return h


def test_systolic_oracle_matches_python_restatement():
    seed = 0x52544C4600000004
    got = oracle.systolic_run(seed, 2, 600)
    for s in range(2):
        assert int(got[s]) == _py_systolic(seed, s, 600)


def test_rv32i_pipeline_model_matches_iss():
    """The cycle-accurate pipeline model must reach the architectural state
    the instruction-set simulator computes (forwarding, stalls, flushes)."""
    import ctypes as C
    from rtlflow_amd.designs import rv32i
    L = oracle.lib()
    L.or_rv32i_state.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    L.or_rv32i_iss.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    progs = rv32i.programs(0x52544C4600000003, np.arange(24))
    for p in progs:
        p = np.ascontiguousarray(p)
        rf1, rf2 = np.zeros(32, np.uint32), np.zeros(32, np.uint32)
        dm1, dm2 = np.zeros(4096, np.uint32), np.zeros(4096, np.uint32)
        retired = L.or_rv32i_iss(p.ctypes.data, 4096, 4095 * 4, 100000, rf1.ctypes.data, dm1.ctypes.data)
        assert 1000 < retired < 4096
        L.or_rv32i_state(p.ctypes.data, 4096, 3 * 4096 + 50, rf2.ctypes.data, dm2.ctypes.data)
        assert (rf1 == rf2).all() and (dm1 == dm2).all()
        assert (rf1 != 0).sum() > 4 and (dm1 != 0).sum() > 10


def test_rv32i_programs_are_valid():
    from rtlflow_amd.designs import rv32i
    p = rv32i.programs(7, np.arange(8))
    assert p.shape == (8, 4096) and (p[:, -1] == 0x63).all()
    ops = p & 0x7f
    assert set(np.unique(ops)) <= {0x33, 0x13, 0x37, 0x17, 0x03, 0x23, 0x63}
    br = p[ops == 0x63]
    assert len(br) > 100
    b = rv32i.budgets(1, 20000)
    assert b.min() >= 1000 and b.max() <= 100000 and 4500 < np.median(b) < 5500
