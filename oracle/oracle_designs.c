/* oracle_designs.c -- golden models of the project's hand-written designs
 * (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Each model is written from the design's specification (the comment block of
 * rtlflow_amd/designs/<design>.py), not from the Verilog: registers are
 * updated from their previous values at the rising edge (non-blocking
 * semantics), inputs for cycle c are r(seed, s, c, port << 3), outputs are
 * hashed after the rising-edge evaluation.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint64_t fnv_u32(uint64_t h, uint32_t v) {
    uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
    return or_fnv(h, b, 4);
}

/* ------------------------------------------------------------------ ALU */

typedef struct {
    uint32_t rf[32];
    uint32_t s1_op, s1_rs1, s1_rs2, s1_rd, s1_imm;
    uint32_t s2_op, s2_rd, s2_a, s2_b;
    uint32_t s3_rd, s3_res;
    uint32_t s4_res;
} alu_t;

static uint32_t alu_exec(uint32_t op, uint32_t a, uint32_t b) {
    switch (op & 7u) {
        case 0: return a + b;
        case 1: return a - b;
        case 2: return a & b;
        case 3: return a | b;
        case 4: return a ^ b;
        case 5: return a << (b & 31u);
        case 6: return a >> (b & 31u);
        default: return a * b;
    }
}

static uint64_t alu_one(uint64_t seed, uint64_t s, uint32_t cycles) {
    alu_t x;
    memset(&x, 0, sizeof x);
    uint64_t h = OR_FNV_BASIS;
    for (uint32_t c = 0; c < cycles; ++c) {
        const uint32_t op = (uint32_t)or_stim(seed, s, c, 0u << 3) & 7u;
        const uint32_t rs1 = (uint32_t)or_stim(seed, s, c, 1u << 3) & 31u;
        const uint32_t rs2 = (uint32_t)or_stim(seed, s, c, 2u << 3) & 31u;
        const uint32_t rd = (uint32_t)or_stim(seed, s, c, 3u << 3) & 31u;
        const uint32_t imm = (uint32_t)or_stim(seed, s, c, 4u << 3);
        alu_t n = x;  /* next state from the current one */
        n.s1_op = op;
        n.s1_rs1 = rs1;
        n.s1_rs2 = rs2;
        n.s1_rd = rd;
        n.s1_imm = imm;
        n.s2_op = x.s1_op;
        n.s2_rd = x.s1_rd;
        n.s2_a = x.rf[x.s1_rs1];
        n.s2_b = x.rf[x.s1_rs2] ^ x.s1_imm;
        n.s3_rd = x.s2_rd;
        n.s3_res = alu_exec(x.s2_op, x.s2_a, x.s2_b);
        n.rf[x.s3_rd] = x.s3_res;
        n.s4_res = x.s3_res;
        x = n;
        h = fnv_u32(h, x.s4_res);
    }
    return h;
}

int or_alu_run(uint64_t seed, uint32_t stim0, uint32_t nstim, uint32_t cycles, uint64_t* hashes,
               int threads) {
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 8)
    for (int64_t i = 0; i < (int64_t)nstim; ++i)
        hashes[i] = alu_one(seed, (uint64_t)stim0 + (uint64_t)i, cycles);
    return 0;
}

/* ------------------------------------------------------------- systolic */

#define SN 16
#define CHAIN_WORDS 128 /* 4,096 bits */

typedef struct {
    int8_t a[SN][SN], b[SN][SN];
    uint32_t acc[SN][SN];
    uint32_t chain[CHAIN_WORDS];
    uint32_t cmp, par;
} sys_t;

static uint64_t systolic_one(uint64_t seed, uint64_t s, uint32_t cycles, sys_t* x, sys_t* n) {
    memset(x, 0, sizeof *x);
    uint64_t h = OR_FNV_BASIS;
    for (uint32_t c = 0; c < cycles; ++c) {
        uint8_t row[SN], col[SN];
        for (int k = 0; k < 2; ++k) {
            const uint64_t r = or_stim(seed, s, c, (0u << 3) | (uint32_t)k);
            const uint64_t q = or_stim(seed, s, c, (1u << 3) | (uint32_t)k);
            for (int i = 0; i < 8; ++i) {
                row[8 * k + i] = (uint8_t)(r >> (8 * i));
                col[8 * k + i] = (uint8_t)(q >> (8 * i));
            }
        }
        for (int i = 0; i < SN; ++i)
            for (int j = 0; j < SN; ++j) {
                n->a[i][j] = j == 0 ? (int8_t)row[i] : x->a[i][j - 1];
                n->b[i][j] = i == 0 ? (int8_t)col[j] : x->b[i - 1][j];
                n->acc[i][j] = x->acc[i][j] + (uint32_t)((int32_t)x->a[i][j] * (int32_t)x->b[i][j]);
            }
        /* chain << 8 | (row0 ^ col0) */
        for (int w = CHAIN_WORDS - 1; w > 0; --w)
            n->chain[w] = (x->chain[w] << 8) | (x->chain[w - 1] >> 24);
        n->chain[0] = (x->chain[0] << 8) | (uint32_t)(row[0] ^ col[0]);
        /* cmp: upper 2048 bits < lower 2048 bits (unsigned), from the old chain */
        int lt = 0;
        for (int w = CHAIN_WORDS / 2 - 1; w >= 0; --w) {
            const uint32_t hi = x->chain[w + CHAIN_WORDS / 2], lo = x->chain[w];
            if (hi != lo) {
                lt = hi < lo;
                break;
            }
        }
        n->cmp = (uint32_t)lt;
        uint32_t p = 0;
        for (int w = 0; w < CHAIN_WORDS; ++w) p ^= x->chain[w];
        n->par = (uint32_t)__builtin_popcount(p) & 1u;
        *x = *n;
        uint32_t ax = 0;
        for (int i = 0; i < SN; ++i)
            for (int j = 0; j < SN; ++j) ax ^= x->acc[i][j];
        h = fnv_u32(h, ax);
        const uint8_t fl = (uint8_t)((x->cmp << 1) | x->par);
        h = or_fnv(h, &fl, 1);
    }
    return h;
}

int or_systolic_run(uint64_t seed, uint32_t stim0, uint32_t nstim, uint32_t cycles,
                    uint64_t* hashes, int threads) {
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
    {
        sys_t* x = (sys_t*)malloc(sizeof(sys_t));
        sys_t* n = (sys_t*)malloc(sizeof(sys_t));
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < (int64_t)nstim; ++i)
            hashes[i] = systolic_one(seed, (uint64_t)stim0 + (uint64_t)i, cycles, x, n);
        free(x);
        free(n);
    }
    return 0;
}

/* --------------------------------------------------------------- RV32I */

enum {
    RV_OP = 0x33, RV_OPI = 0x13, RV_LUI = 0x37, RV_AUIPC = 0x17, RV_LOAD = 0x03,
    RV_STORE = 0x23, RV_BRANCH = 0x63, RV_JAL = 0x6f, RV_JALR = 0x67
};

typedef struct {
    uint32_t dmem[4096];
    uint32_t rf[32];
    uint32_t pc;
    uint32_t d_v, d_pc, d_ir;
    uint32_t e_v, e_wen, e_f7, e_pc, e_a, e_b, e_imm, e_rd, e_rs1, e_rs2, e_op, e_f3;
    uint32_t m_v, m_wen, m_res, m_sd, m_rd, m_op, m_f3;
    uint32_t w_wen, w_res, w_rd;
} rv_t;

static uint32_t rv_sx(uint32_t v, int bits) {
    return (uint32_t)(((int32_t)(v << (32 - bits))) >> (32 - bits));
}

static uint32_t rv_imm(uint32_t ir) {
    switch (ir & 0x7f) {
        case RV_STORE: return rv_sx(((ir >> 25) << 5) | ((ir >> 7) & 31), 12);
        case RV_BRANCH:
            return rv_sx((((ir >> 31) & 1) << 12) | (((ir >> 7) & 1) << 11) | (((ir >> 25) & 0x3f) << 5) |
                             (((ir >> 8) & 0xf) << 1), 13);
        case RV_LUI:
        case RV_AUIPC: return ir & 0xfffff000u;
        case RV_JAL:
            return rv_sx((((ir >> 31) & 1) << 20) | (((ir >> 12) & 0xff) << 12) | (((ir >> 20) & 1) << 11) |
                             (((ir >> 21) & 0x3ff) << 1), 21);
        default: return rv_sx(ir >> 20, 12);
    }
}

static int rv_writes(uint32_t op) {
    return op == RV_OP || op == RV_OPI || op == RV_LUI || op == RV_AUIPC || op == RV_LOAD ||
           op == RV_JAL || op == RV_JALR;
}

static uint32_t rv_alu(uint32_t op, uint32_t f3, uint32_t f7, uint32_t a, uint32_t b) {
    switch (f3) {
        case 0: return (op == RV_OP && f7) ? a - b : a + b;
        case 1: return a << (b & 31);
        case 2: return (int32_t)a < (int32_t)b;
        case 3: return a < b;
        case 4: return a ^ b;
        case 5: return f7 ? (uint32_t)((int32_t)a >> (b & 31)) : a >> (b & 31);
        case 6: return a | b;
        default: return a & b;
    }
}

static int rv_branch(uint32_t f3, uint32_t a, uint32_t b) {
    switch (f3) {
        case 0: return a == b;
        case 1: return a != b;
        case 4: return (int32_t)a < (int32_t)b;
        case 5: return (int32_t)a >= (int32_t)b;
        case 6: return a < b;
        case 7: return a >= b;
        default: return 0;
    }
}

static uint32_t rv_load(uint32_t f3, uint32_t word, uint32_t addr) {
    const uint32_t byte = (word >> ((addr & 3) * 8)) & 0xff;
    const uint32_t half = (addr & 2) ? (word >> 16) : (word & 0xffff);
    switch (f3) {
        case 0: return rv_sx(byte, 8);
        case 1: return rv_sx(half, 16);
        case 4: return byte;
        case 5: return half;
        default: return word;
    }
}

static uint32_t rv_merge(uint32_t f3, uint32_t word, uint32_t addr, uint32_t data) {
    switch (f3 & 3) {
        case 0: {
            const uint32_t sh = (addr & 3) * 8;
            return (word & ~(0xffu << sh)) | ((data & 0xff) << sh);
        }
        case 1: return (addr & 2) ? ((word & 0xffff) | (data << 16)) : ((word & 0xffff0000u) | (data & 0xffff));
        default: return data;
    }
}

/* One rising edge of the pipeline (all next values from the current state). */
static void rv_step(rv_t* x, const uint32_t* prog, uint32_t nwords) {
    /* ID */
    const uint32_t d_op = x->d_ir & 0x7f, d_rd = (x->d_ir >> 7) & 31, d_f3 = (x->d_ir >> 12) & 7;
    const uint32_t d_rs1 = (x->d_ir >> 15) & 31, d_rs2 = (x->d_ir >> 20) & 31;
    const uint32_t d_imm = rv_imm(x->d_ir);
    const uint32_t d_wen = x->d_v && rv_writes(d_op) && d_rd != 0;
    const uint32_t d_a = (x->w_wen && x->w_rd == d_rs1) ? x->w_res : x->rf[d_rs1];
    const uint32_t d_b = (x->w_wen && x->w_rd == d_rs2) ? x->w_res : x->rf[d_rs2];
    /* EX */
    const uint32_t x_a = (x->m_wen && x->m_op != RV_LOAD && x->m_rd == x->e_rs1) ? x->m_res
                         : (x->w_wen && x->w_rd == x->e_rs1) ? x->w_res : x->e_a;
    const uint32_t x_b = (x->m_wen && x->m_op != RV_LOAD && x->m_rd == x->e_rs2) ? x->m_res
                         : (x->w_wen && x->w_rd == x->e_rs2) ? x->w_res : x->e_b;
    const uint32_t x_op2 = x->e_op == RV_OP ? x_b : x->e_imm;
    uint32_t x_res = 0, x_target = x->e_pc + x->e_imm;
    int x_taken = 0;
    switch (x->e_op) {
        case RV_OP:
        case RV_OPI: x_res = rv_alu(x->e_op, x->e_f3, x->e_f7, x_a, x_op2); break;
        case RV_LUI: x_res = x->e_imm; break;
        case RV_AUIPC: x_res = x->e_pc + x->e_imm; break;
        case RV_LOAD:
        case RV_STORE: x_res = x_a + x->e_imm; break;
        case RV_BRANCH: x_taken = rv_branch(x->e_f3, x_a, x_b); break;
        case RV_JAL: x_res = x->e_pc + 4; x_taken = 1; break;
        case RV_JALR: x_res = x->e_pc + 4; x_taken = 1; x_target = (x_a + x->e_imm) & ~1u; break;
        default: break;
    }
    if (!x->e_v) x_taken = 0;
    const int load_use = x->e_v && x->e_op == RV_LOAD && x->e_wen && x->d_v &&
                         (x->e_rd == d_rs1 || x->e_rd == d_rs2);
    /* MEM */
    const uint32_t widx = (x->m_res >> 2) & 4095;
    const uint32_t m_word = x->dmem[widx];
    const uint32_t m_load = rv_load(x->m_f3, m_word, x->m_res);
    const uint32_t m_merged = rv_merge(x->m_f3, m_word, x->m_res, x->m_sd);
    /* commit */
    rv_t n = *x;
    if (x->w_wen) n.rf[x->w_rd] = x->w_res;
    n.w_wen = x->m_wen;
    n.w_rd = x->m_rd;
    n.w_res = x->m_op == RV_LOAD ? m_load : x->m_res;
    if (x->m_v && x->m_op == RV_STORE) n.dmem[widx] = m_merged;
    n.m_v = x->e_v;
    n.m_wen = x->e_v && x->e_wen;
    n.m_res = x_res;
    n.m_sd = x_b;
    n.m_rd = x->e_rd;
    n.m_op = x->e_op;
    n.m_f3 = x->e_f3;
    if (x_taken || load_use) {
        n.e_v = 0;
        n.e_wen = 0;
        n.e_op = 0;
    } else {
        n.e_v = x->d_v;
        n.e_wen = d_wen;
        n.e_pc = x->d_pc;
        n.e_a = d_a;
        n.e_b = d_b;
        n.e_imm = d_imm;
        n.e_rd = d_rd;
        n.e_rs1 = d_rs1;
        n.e_rs2 = d_rs2;
        n.e_op = d_op;
        n.e_f3 = d_f3;
        n.e_f7 = (x->d_ir >> 30) & 1;
    }
    if (x_taken) {
        n.d_v = 0;
    } else if (!load_use) {
        const uint32_t ia = (x->pc >> 2) & 16383;
        n.d_v = 1;
        n.d_pc = x->pc;
        n.d_ir = ia < nwords ? prog[ia] : 0;
    }
    if (x_taken)
        n.pc = x_target;
    else if (!load_use)
        n.pc = x->pc + 4;
    *x = n;
}

static uint64_t rv_one(const uint32_t* prog, uint32_t nwords, uint32_t cycles, rv_t* x) {
    memset(x, 0, sizeof *x);
    uint64_t h = OR_FNV_BASIS;
    for (uint32_t c = 0; c < cycles; ++c) {
        rv_step(x, prog, nwords);
        h = fnv_u32(h, x->pc);
        h = fnv_u32(h, x->w_wen ? x->w_res : 0u);
    }
    return h;
}

int or_rv32i_run(const uint32_t* programs, uint32_t nwords, uint64_t seed, uint32_t stim0,
                 uint32_t nstim, const uint32_t* budgets, uint64_t* hashes, int threads) {
    (void)seed;
    (void)stim0;
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
    {
        rv_t* x = (rv_t*)malloc(sizeof(rv_t));
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < (int64_t)nstim; ++i)
            hashes[i] = rv_one(programs + (size_t)i * nwords, nwords, budgets[i], x);
        free(x);
    }
    return 0;
}

/* Pipeline state after ``cycles`` edges (register file and data RAM), for
 * checking the pipeline model against the instruction-set simulator. */
int or_rv32i_state(const uint32_t* prog, uint32_t nwords, uint32_t cycles, uint32_t* rf_out,
                   uint32_t* dmem_out) {
    rv_t* x = (rv_t*)malloc(sizeof(rv_t));
    memset(x, 0, sizeof *x);
    for (uint32_t c = 0; c < cycles; ++c) rv_step(x, prog, nwords);
    memcpy(rf_out, x->rf, sizeof x->rf);
    memcpy(dmem_out, x->dmem, sizeof x->dmem);
    free(x);
    return 0;
}

/* Instruction-set simulator: architectural execution from pc 0 until the
 * self-loop at ``halt_pc`` (or max_steps). Returns the number of retired
 * instructions. Timing-free: pins what the pipeline must compute. */
int or_rv32i_iss(const uint32_t* prog, uint32_t nwords, uint32_t halt_pc, uint32_t max_steps,
                 uint32_t* rf_out, uint32_t* dmem_out) {
    uint32_t rf[32] = {0};
    uint32_t* dm = (uint32_t*)calloc(4096, 4);
    uint32_t pc = 0;
    int n = 0;
    for (; (uint32_t)n < max_steps && pc != halt_pc; ++n) {
        const uint32_t ia = (pc >> 2) & 16383;
        const uint32_t ir = ia < nwords ? prog[ia] : 0;
        const uint32_t op = ir & 0x7f, rd = (ir >> 7) & 31, f3 = (ir >> 12) & 7;
        const uint32_t a = rf[(ir >> 15) & 31], b = rf[(ir >> 20) & 31], imm = rv_imm(ir);
        uint32_t npc = pc + 4, val = 0;
        int wr = 0;
        switch (op) {
            case RV_OP: val = rv_alu(op, f3, (ir >> 30) & 1, a, b); wr = 1; break;
            case RV_OPI: val = rv_alu(op, f3, (ir >> 30) & 1, a, imm); wr = 1; break;
            case RV_LUI: val = imm; wr = 1; break;
            case RV_AUIPC: val = pc + imm; wr = 1; break;
            case RV_LOAD: {
                const uint32_t ad = a + imm;
                val = rv_load(f3, dm[(ad >> 2) & 4095], ad);
                wr = 1;
                break;
            }
            case RV_STORE: {
                const uint32_t ad = a + imm;
                dm[(ad >> 2) & 4095] = rv_merge(f3, dm[(ad >> 2) & 4095], ad, b);
                break;
            }
            case RV_BRANCH:
                if (rv_branch(f3, a, b)) npc = pc + imm;
                break;
            case RV_JAL: val = pc + 4; wr = 1; npc = pc + imm; break;
            case RV_JALR: val = pc + 4; wr = 1; npc = (a + imm) & ~1u; break;
            default: break;
        }
        if (wr && rd) rf[rd] = val;
        pc = npc;
    }
    memcpy(rf_out, rf, sizeof rf);
    memcpy(dmem_out, dm, 4096 * 4);
    free(dm);
    return n;
}
