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

int or_systolic_run(uint64_t seed, uint32_t stim0, uint32_t nstim, uint32_t cycles,
                    uint64_t* hashes, int threads) {
    (void)seed; (void)stim0; (void)nstim; (void)cycles; (void)hashes; (void)threads;
    return -1;
}

int or_rv32i_run(const uint32_t* programs, uint32_t nwords, uint64_t seed, uint32_t stim0,
                 uint32_t nstim, const uint32_t* budgets, uint64_t* hashes, int threads) {
    (void)programs; (void)nwords; (void)seed; (void)stim0; (void)nstim; (void)budgets;
    (void)hashes; (void)threads;
    return -1;
}
