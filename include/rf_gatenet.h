/* rf_gatenet.h -- gate-level netlist engine for designs beyond the per-design
 * emitter's size (part of librf_verilated.so).
 *
 * A structural netlist of single-bit gates (AND, OR, XOR, NAND, 2:1 MUX), D
 * flip-flops on one clock and synchronous RAMs (registered read port, one
 * write port) runs the benchmark's clock-cycle protocol for one stimulus per
 * lane: per cycle, the primary inputs from the stimulus generator with
 * clk = 0 and an eval, the rising edge (flip-flops and RAM ports take the
 * settled values), an eval with clk = 1, then the primary outputs folded into
 * the lane's FNV-1a-64 hash while the cycle is below the lane's budget.
 *
 * The netlist arrives as a program blob written by rtlflow_amd.gatenet: the
 * gates in levels, the flip-flop, output and RAM port tables, and a
 * partition of the gates into mtasks (a DAG) from the V3Partition
 * restatement. Two forms:
 *
 *   RF_GN_GRAPH     the plain per-mtask form: one lane per stimulus, a byte
 *                   per node in a lane-minor struct of arrays, one kernel per
 *                   mtask per eval, replayed as a hipGraph per clock cycle;
 *   RF_GN_BITPLANE  64 stimuli per 64-bit plane, planes of all 64-lane
 *                   groups of one node contiguous (one coalesced access per
 *                   wavefront for 64 groups), one kernel per level, the
 *                   clk = 1 eval limited to the outputs' cone. Stimuli are
 *                   dealt to lanes longest budget first, and each cycle runs
 *                   over the first power-of-two count of groups that covers
 *                   every group with a lane still running (one graph per
 *                   count), so finished groups cost nothing.
 *
 * No counterpart in the reference, which compiles every design through its
 * emitter (src/V3EmitC.cpp); the protocol and hash are the benchmark's.
 * Errors as rf_verilated.h: RF_OK or a negative RF_E_* code, message from
 * rf_last_error(). */
#ifndef RF_GATENET_H
#define RF_GATENET_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { RF_GN_GRAPH = 0, RF_GN_BITPLANE = 1 };

typedef struct rf_gatenet rf_gatenet;

/* Check and load a program for n_lanes stimuli (stimulus id = lane). */
int rf_gn_create(const uint32_t* program, size_t words, uint32_t n_lanes, rf_gatenet** out);

/* Run every lane from the reset state for budgets[lane] cycles (the launch
 * runs the largest budget; a lane stops hashing at its own) in ``mode``;
 * out_hashes[n_lanes] receives the hashes, *ms (may be NULL) the simulation
 * time measured with HIP events on the engine's stream. */
int rf_gn_run(rf_gatenet* g, uint32_t mode, uint64_t seed, const uint32_t* budgets,
              uint64_t* out_hashes, float* ms);

/* Mtasks of the loaded partition, graph nodes per clock cycle of a mode. */
uint32_t rf_gn_mtasks(const rf_gatenet* g);
uint32_t rf_gn_nodes_per_cycle(const rf_gatenet* g, uint32_t mode);

int rf_gn_destroy(rf_gatenet* g);

#ifdef __cplusplus
}
#endif
#endif
