/* oracle.h -- CPU oracle for parity tests (TEST INFRASTRUCTURE ONLY).
 *
 * Plain-C restatement of what the simulated designs compute, written
 * independently of the rtlflow_amd compiler: it never sees the generated
 * code or the compiler's IR. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or the
 * timed CPU baseline.
 *
 * Protocol restated (identical on CPU and GPU): per stimulus s and cycle c,
 * apply inputs r(seed, s, c, k) with clk = 0 and evaluate, set clk = 1 and
 * evaluate (flip-flops take the values settled in the clk = 0 evaluation),
 * then fold the output port's little-endian bytes into an FNV-1a-64 hash.
 * State starts at zero (Verilator --x-assign 0 --x-initial 0).
 */
#ifndef RF_ORACLE_H
#define RF_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_FNV_BASIS 0xcbf29ce484222325ull
#define OR_FNV_PRIME 0x100000001b3ull

uint64_t or_splitmix64(uint64_t x);
uint64_t or_stim(uint64_t seed, uint64_t s, uint64_t c, uint64_t k);
uint64_t or_fnv(uint64_t h, const uint8_t* bytes, size_t n);

/* Random gate netlist (binary gate table written by the netlist generator).
 * budgets: cycles per stimulus (NULL: all ``cycles``). Returns 0 on success. */
int or_netlist_run(const uint8_t* table, size_t len, uint64_t seed, uint32_t stim0,
                   uint32_t nstim, uint32_t cycles, const uint32_t* budgets, uint64_t* hashes,
                   int threads);

/* 4-stage pipelined 32-bit ALU with a 32x32 register file. */
int or_alu_run(uint64_t seed, uint32_t stim0, uint32_t nstim, uint32_t cycles, uint64_t* hashes,
               int threads);

/* 16x16 int8 systolic array + 4096-bit shift/compare chain. */
int or_systolic_run(uint64_t seed, uint32_t stim0, uint32_t nstim, uint32_t cycles,
                    uint64_t* hashes, int threads);

/* 5-stage RV32I pipeline; programs[s] = 4096-instruction image of stimulus s
 * (nwords words each), budgets[s] cycles. */
int or_rv32i_run(const uint32_t* programs, uint32_t nwords, uint64_t seed, uint32_t stim0,
                 uint32_t nstim, const uint32_t* budgets, uint64_t* hashes, int threads);

#ifdef __cplusplus
}
#endif
#endif
