/* oracle_core.c -- stimulus generator, output hash and the random-netlist
 * evaluator of the CPU oracle (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * SplitMix64: the published finalizer (Steele, Lea, Flood, OOPSLA 2014).
 * FNV-1a 64: offset basis 0xcbf29ce484222325, prime 0x100000001b3 (Fowler,
 * Noll, Vo). Known answers are pinned in tests/test_oracle.py.
 *
 * Netlist semantics restated from the generator's definition (the oracle reads
 * the gate table, never the Verilog): node values are bits; AND/OR/XOR/NAND
 * of two fan-ins, MUX = sel ? a : b. A flip-flop takes, at the rising edge,
 * the value its D node had after the clk = 0 evaluation. A RAM writes
 * wdata at addr when we is set, and its registered read port takes the word
 * at addr from before the write (non-blocking semantics), at the same edge.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

uint64_t or_splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t or_stim(uint64_t seed, uint64_t s, uint64_t c, uint64_t k) {
    return or_splitmix64(seed ^ (s * 0x9E3779B97F4A7C15ull) ^ (c << 8) ^ k);
}

uint64_t or_fnv(uint64_t h, const uint8_t* bytes, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        h ^= bytes[i];
        h *= OR_FNV_PRIME;
    }
    return h;
}

static uint64_t fnv_le(uint64_t h, uint64_t v, int nbytes) {
    uint8_t b[8];
    for (int i = 0; i < nbytes; ++i) b[i] = (uint8_t)(v >> (8 * i));
    return or_fnv(h, b, (size_t)nbytes);
}

typedef struct {
    uint32_t n_pi, n_ff, n_ram_bits, n_gates, n_po, n_rams, ram_depth, ram_width, abits;
    const uint32_t* type;
    const uint32_t* in;   /* [n_gates][3] */
    const uint32_t* ff_d;
    const uint32_t* po;
    const uint32_t* ram;  /* per RAM: addr[abits], we, wdata[ram_width] */
} table_t;

static int parse_table(const uint8_t* buf, size_t len, table_t* t) {
    if (len < 40 || memcmp(buf, "RFNL", 4) != 0) return -1;
    const uint32_t* h = (const uint32_t*)(buf + 4);
    t->n_pi = h[0];
    t->n_ff = h[1];
    t->n_ram_bits = h[2];
    t->n_gates = h[3];
    t->n_po = h[4];
    t->n_rams = h[5];
    t->ram_depth = h[6];
    t->ram_width = h[7];
    t->abits = h[8];
    const uint32_t* p = h + 9;
    t->type = p;
    p += t->n_gates;
    t->in = p;
    p += 3ull * t->n_gates;
    t->ff_d = p;
    p += t->n_ff;
    t->po = p;
    p += t->n_po;
    t->ram = p;
    p += (size_t)t->n_rams * (t->abits + 1 + t->ram_width);
    if ((const uint8_t*)p > buf + len) return -2;
    if (t->n_pi > 128 || t->n_po > 64 || (t->n_rams && t->ram_width > 64)) return -3;
    return 0;
}

static void eval_comb(const table_t* t, uint8_t* v) {
    const uint32_t first = t->n_pi + t->n_ff + t->n_ram_bits;
    for (uint32_t g = 0; g < t->n_gates; ++g) {
        const uint32_t* in = t->in + 3ull * g;
        const uint8_t a = v[in[0]], b = v[in[1]];
        uint8_t r;
        switch (t->type[g]) {
            case 0: r = a & b; break;
            case 1: r = a | b; break;
            case 2: r = a ^ b; break;
            case 3: r = (uint8_t)((a & b) ^ 1u); break;
            default: r = a ? b : v[in[2]]; break;
        }
        v[first + g] = r;
    }
}

static uint64_t run_one(const table_t* t, uint64_t seed, uint64_t s, uint32_t cycles, uint8_t* v,
                        uint64_t* ram, uint8_t* dtmp) {
    const uint32_t nodes = t->n_pi + t->n_ff + t->n_ram_bits + t->n_gates;
    memset(v, 0, nodes);
    if (t->n_rams) memset(ram, 0, (size_t)t->n_rams * t->ram_depth * 8);
    /* primary inputs: bits [64j, 64j + 64) of ``pi`` from draw k = j */
    const uint32_t npiw = (t->n_pi + 63) / 64;
    const uint32_t obytes = t->n_po <= 8 ? 1 : t->n_po <= 16 ? 2 : t->n_po <= 32 ? 4 : 8;
    uint64_t h = OR_FNV_BASIS;
    const uint32_t rd0 = t->n_pi + t->n_ff;
    const uint64_t wmask = t->ram_width >= 64 ? ~0ull : ((1ull << t->ram_width) - 1);
    for (uint32_t c = 0; c < cycles; ++c) {
        for (uint32_t j = 0; j < npiw; ++j) {
            const uint64_t pi = or_stim(seed, s, c, j);
            for (uint32_t i = 64 * j; i < t->n_pi && i < 64 * j + 64; ++i)
                v[i] = (uint8_t)((pi >> (i - 64 * j)) & 1u);
        }
        eval_comb(t, v);                                   /* clk = 0 eval */
        for (uint32_t i = 0; i < t->n_ff; ++i) dtmp[i] = v[t->ff_d[i]];
        for (uint32_t r = 0; r < t->n_rams; ++r) {         /* posedge: RAMs */
            const uint32_t* rp = t->ram + (size_t)r * (t->abits + 1 + t->ram_width);
            uint64_t addr = 0, wd = 0;
            for (uint32_t k = 0; k < t->abits; ++k) addr |= (uint64_t)v[rp[k]] << k;
            const uint8_t we = v[rp[t->abits]];
            for (uint32_t k = 0; k < t->ram_width; ++k) wd |= (uint64_t)v[rp[t->abits + 1 + k]] << k;
            uint64_t* mem = ram + (size_t)r * t->ram_depth;
            const uint64_t old = addr < t->ram_depth ? mem[addr] : 0;
            if (we && addr < t->ram_depth) mem[addr] = wd & wmask;
            for (uint32_t k = 0; k < t->ram_width; ++k)
                v[rd0 + r * t->ram_width + k] = (uint8_t)((old >> k) & 1u);
        }
        for (uint32_t i = 0; i < t->n_ff; ++i) v[t->n_pi + i] = dtmp[i];  /* posedge: flops */
        eval_comb(t, v);                                   /* clk = 1 eval */
        uint64_t po = 0;
        for (uint32_t k = 0; k < t->n_po; ++k) po |= (uint64_t)v[t->po[k]] << k;
        h = fnv_le(h, po, (int)obytes);
    }
    return h;
}

int or_netlist_run(const uint8_t* table, size_t len, uint64_t seed, uint32_t stim0,
                   uint32_t nstim, uint32_t cycles, const uint32_t* budgets, uint64_t* hashes,
                   int threads) {
    table_t t;
    int rc = parse_table(table, len, &t);
    if (rc) return rc;
    const uint32_t nodes = t.n_pi + t.n_ff + t.n_ram_bits + t.n_gates;
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
    {
        uint8_t* v = (uint8_t*)malloc(nodes);
        uint8_t* dtmp = (uint8_t*)malloc(t.n_ff + 1);
        uint64_t* ram = (uint64_t*)malloc((size_t)t.n_rams * t.ram_depth * 8 + 8);
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < (int64_t)nstim; ++i) {
            const uint32_t cyc = budgets ? budgets[i] : cycles;
            hashes[i] = run_one(&t, seed, (uint64_t)stim0 + (uint64_t)i, cyc, v, ram, dtmp);
        }
        free(v);
        free(dtmp);
        free(ram);
    }
    return 0;
}
