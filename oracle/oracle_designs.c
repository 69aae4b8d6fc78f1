/* oracle_designs.c -- golden models of the project's hand-written designs
 * (TEST INFRASTRUCTURE ONLY; see oracle.h). */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int or_alu_run(uint64_t seed, uint32_t stim0, uint32_t nstim, uint32_t cycles, uint64_t* hashes,
               int threads) {
    (void)seed; (void)stim0; (void)nstim; (void)cycles; (void)hashes; (void)threads;
    return -1;
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
